/**
 * @file spm.h
 *
 * The sparse matrix descriptor consumed by the PaStiX C API.
 *
 * The struct layout and enum values are those of the reference SPM library
 * (spm/include/spm.h:50-78, spm/include/spm_const.h, spm_datatypes.h:25-51),
 * because pastix_task_analyze/numfact take a spmatrix_t* built by the caller.
 * Only the type is provided here: callers keep using their own SPM library to
 * build and manipulate matrices.
 */
#ifndef PASTIX_AMD_SPM_H
#define PASTIX_AMD_SPM_H

#include <stdint.h>

#if defined(SPM_INT32)
typedef int32_t spm_int_t;
#else
typedef int64_t spm_int_t;
#endif

typedef enum spm_coeftype_e {
    SpmPattern   = 0,
    SpmFloat     = 2,
    SpmDouble    = 3,
    SpmComplex32 = 4,
    SpmComplex64 = 5
} spm_coeftype_t;

typedef enum spm_fmttype_e {
    SpmCSC,
    SpmCSR,
    SpmIJV
} spm_fmttype_t;

typedef enum spm_layout_e {
    SpmRowMajor = 101,
    SpmColMajor = 102
} spm_layout_t;

typedef enum spm_trans_e {
    SpmNoTrans   = 111,
    SpmTrans     = 112,
    SpmConjTrans = 113
} spm_trans_t;

typedef enum spm_mtxtype_e {
    SpmGeneral   = SpmNoTrans,
    SpmSymmetric = SpmTrans,
    SpmHermitian = SpmConjTrans
} spm_mtxtype_t;

typedef enum spm_normtype_e {
    SpmOneNorm       = 171,
    SpmFrobeniusNorm = 174,
    SpmInfNorm       = 175,
    SpmMaxNorm       = 177
} spm_normtype_t;

typedef enum spm_driver_e {
    SpmDriverRSA,
    SpmDriverHB,
    SpmDriverIJV,
    SpmDriverMM,
    SpmDriverLaplacian,
    SpmDriverXLaplacian,
    SpmDriverGraph,
    SpmDriverSPM
} spm_driver_t;

typedef struct spmatrix_s {
    spm_mtxtype_t  mtxtype;
    spm_coeftype_t flttype;
    spm_fmttype_t  fmttype;

    spm_int_t      gN;
    spm_int_t      n;
    spm_int_t      gnnz;
    spm_int_t      nnz;

    spm_int_t      gNexp;
    spm_int_t      nexp;
    spm_int_t      gnnzexp;
    spm_int_t      nnzexp;

    spm_int_t      dof;
    spm_int_t     *dofs;
    spm_layout_t   layout;

    spm_int_t     *colptr;
    spm_int_t     *rowptr;
    spm_int_t     *loc2glob;
    void          *values;
} spmatrix_t;

#endif
