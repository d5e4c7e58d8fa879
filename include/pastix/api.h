/**
 * @file pastix/api.h
 *
 * Parameter indices and enumerations of the PaStiX C API.
 *
 * The numeric values are part of the ABI: iparm[]/dparm[] are arrays indexed
 * by these enums in caller code compiled against the reference headers.  The
 * order and values follow the reference include/pastix/api.h:26-160 (iparm,
 * dparm) and :168-468 (tasks, factorization, scheduler, refinement, errors,
 * BLAS-like enums).  Nothing is renumbered; engine-specific knobs live in
 * pastix_amd.h in a separate array.
 */
#ifndef PASTIX_AMD_API_H
#define PASTIX_AMD_API_H

typedef enum pastix_iparm_e {
    IPARM_VERBOSE,
    IPARM_IO_STRATEGY,
    IPARM_NNZEROS,
    IPARM_NNZEROS_BLOCK_LOCAL,
    IPARM_ALLOCATED_TERMS,
    IPARM_PRODUCE_STATS,
    IPARM_MC64,
    IPARM_ORDERING,
    IPARM_ORDERING_DEFAULT,
    IPARM_SCOTCH_SWITCH_LEVEL,
    IPARM_SCOTCH_CMIN,
    IPARM_SCOTCH_CMAX,
    IPARM_SCOTCH_FRAT,
    IPARM_METIS_CTYPE,
    IPARM_METIS_RTYPE,
    IPARM_METIS_NO2HOP,
    IPARM_METIS_NSEPS,
    IPARM_METIS_NITER,
    IPARM_METIS_UFACTOR,
    IPARM_METIS_COMPRESS,
    IPARM_METIS_CCORDER,
    IPARM_METIS_PFACTOR,
    IPARM_METIS_SEED,
    IPARM_METIS_DBGLVL,
    IPARM_AMALGAMATION_LVLBLAS,
    IPARM_AMALGAMATION_LVLCBLK,
    IPARM_REORDERING_SPLIT,
    IPARM_REORDERING_STOP,
    IPARM_SPLITTING_STRATEGY,
    IPARM_SPLITTING_LEVELS_PROJECTIONS,
    IPARM_SPLITTING_LEVELS_KWAY,
    IPARM_SPLITTING_PROJECTIONS_DEPTH,
    IPARM_SPLITTING_PROJECTIONS_DISTANCE,
    IPARM_SPLITTING_PROJECTIONS_WIDTH,
    IPARM_MIN_BLOCKSIZE,
    IPARM_MAX_BLOCKSIZE,
    IPARM_TASKS2D_LEVEL,
    IPARM_TASKS2D_WIDTH,
    IPARM_ALLCAND,
    IPARM_INCOMPLETE,
    IPARM_LEVEL_OF_FILL,
    IPARM_FACTORIZATION,
    IPARM_STATIC_PIVOTING,
    IPARM_FREE_CSCUSER,
    IPARM_SCHUR_FACT_MODE,
    IPARM_SCHUR_SOLV_MODE,
    IPARM_APPLYPERM_WS,
    IPARM_REFINEMENT,
    IPARM_NBITER,
    IPARM_ITERMAX,
    IPARM_GMRES_IM,
    IPARM_SCHEDULER,
    IPARM_THREAD_NBR,
    IPARM_AUTOSPLIT_COMM,
    IPARM_GPU_NBR,
    IPARM_GPU_MEMORY_PERCENTAGE,
    IPARM_GPU_MEMORY_BLOCK_SIZE,
    IPARM_COMPRESS_MIN_WIDTH,
    IPARM_COMPRESS_MIN_HEIGHT,
    IPARM_COMPRESS_WHEN,
    IPARM_COMPRESS_METHOD,
    IPARM_COMPRESS_ORTHO,
    IPARM_COMPRESS_RELTOL,
    IPARM_COMPRESS_PRESELECT,
    IPARM_THREAD_COMM_MODE,
    IPARM_MODIFY_PARAMETER,
    IPARM_START_TASK,
    IPARM_END_TASK,
    IPARM_FLOAT,
    IPARM_MTX_TYPE,
    IPARM_DOF_NBR,
    IPARM_SIZE
} pastix_iparm_t;

typedef enum pastix_dparm_e {
    DPARM_FILL_IN,
    DPARM_EPSILON_REFINEMENT,
    DPARM_RELATIVE_ERROR,
    DPARM_EPSILON_MAGN_CTRL,
    DPARM_ANALYZE_TIME,
    DPARM_PRED_FACT_TIME,
    DPARM_FACT_TIME,
    DPARM_SOLV_TIME,
    DPARM_FACT_FLOPS,
    DPARM_FACT_THFLOPS,
    DPARM_FACT_RLFLOPS,
    DPARM_SOLV_FLOPS,
    DPARM_SOLV_THFLOPS,
    DPARM_SOLV_RLFLOPS,
    DPARM_REFINE_TIME,
    DPARM_A_NORM,
    DPARM_COMPRESS_TOLERANCE,
    DPARM_COMPRESS_MIN_RATIO,
    DPARM_SIZE
} pastix_dparm_t;

typedef enum pastix_task_e {
    PastixTaskInit     = 0,
    PastixTaskOrdering = 1,
    PastixTaskSymbfact = 2,
    PastixTaskAnalyze  = 3,
    PastixTaskNumfact  = 4,
    PastixTaskSolve    = 5,
    PastixTaskRefine   = 6,
    PastixTaskClean    = 7
} pastix_task_t;

typedef enum pastix_verbose_e {
    PastixVerboseNot = 0,
    PastixVerboseNo  = 1,
    PastixVerboseYes = 2
} pastix_verbose_t;

typedef enum pastix_io_e {
    PastixIONo        = 0,
    PastixIOLoad      = 1,
    PastixIOSave      = 2,
    PastixIOLoadGraph = 4,
    PastixIOSaveGraph = 8,
    PastixIOLoadCSC   = 16,
    PastixIOSaveCSC   = 32
} pastix_io_t;

typedef enum pastix_fact_mode_e {
    PastixFactModeLocal = 0,
    PastixFactModeSchur = 1,
    PastixFactModeBoth  = 2
} pastix_fact_mode_t;

typedef enum pastix_solv_mode_e {
    PastixSolvModeLocal     = 0,
    PastixSolvModeInterface = 1,
    PastixSolvModeSchur     = 2
} pastix_solv_mode_t;

typedef enum pastix_refine_e {
    PastixRefineGMRES,
    PastixRefineCG,
    PastixRefineSR,
    PastixRefineBiCGSTAB
} pastix_refine_t;

#define pastix_coeftype_t spm_coeftype_t
#define PastixPattern   SpmPattern
#define PastixFloat     SpmFloat
#define PastixDouble    SpmDouble
#define PastixComplex32 SpmComplex32
#define PastixComplex64 SpmComplex64

typedef enum pastix_factotype_e {
    PastixFactPOTRF = 0,
    PastixFactSYTRF = 1,
    PastixFactGETRF = 2,
    PastixFactPXTRF = 3,
    PastixFactHETRF = 4,

    PastixFactLLH  = 0,
    PastixFactLDLT = 1,
    PastixFactLU   = 2,
    PastixFactLLT  = 3,
    PastixFactLDLH = 4
} pastix_factotype_t;

typedef enum pastix_scheduler_e {
    PastixSchedSequential = 0,
    PastixSchedStatic     = 1,
    PastixSchedParsec     = 2,
    PastixSchedStarPU     = 3,
    PastixSchedDynamic    = 4
} pastix_scheduler_t;

enum pastix_order_e {
    PastixOrderScotch,
    PastixOrderMetis,
    PastixOrderPersonal,
    PastixOrderPtScotch,
    PastixOrderParMetis
};

/* Threaded communication mode (IPARM_THREAD_COMM_MODE); the reference
 * defines it only in MPI builds (include/pastix/api.h:312-315), values kept. */
typedef enum pastix_threadmode_e {
    PastixThreadMultiple = 1,
    PastixThreadFunneled = 2
} pastix_threadmode_t;

typedef enum pastix_error_e {
    PASTIX_SUCCESS            = 0,
    PASTIX_ERR_UNKNOWN        = 1,
    PASTIX_ERR_ALLOC          = 2,
    PASTIX_ERR_NOTIMPLEMENTED = 3,
    PASTIX_ERR_OUTOFMEMORY    = 4,
    PASTIX_ERR_THREAD         = 5,
    PASTIX_ERR_INTERNAL       = 6,
    PASTIX_ERR_BADPARAMETER   = 7,
    PASTIX_ERR_FILE           = 8,
    PASTIX_ERR_INTEGER_TYPE   = 9,
    PASTIX_ERR_IO             = 10,
    PASTIX_ERR_MPI            = 11
} pastix_error_t;

typedef enum pastix_compress_when_e {
    PastixCompressNever,
    PastixCompressWhenBegin,
    PastixCompressWhenEnd,
    PastixCompressWhenDuring
} pastix_compress_when_t;

typedef enum pastix_compress_method_e {
    PastixCompressMethodSVD,
    PastixCompressMethodPQRCP,
    PastixCompressMethodRQRCP,
    PastixCompressMethodTQRCP,
    PastixCompressMethodRQRRT,
    PastixCompressMethodNbr
} pastix_compress_method_t;

typedef enum pastix_compress_ortho_e {
    PastixCompressOrthoCGS,
    PastixCompressOrthoQR,
    PastixCompressOrthoPartialQR
} pastix_compress_ortho_t;

typedef enum pastix_split_e {
    PastixSplitNot,
    PastixSplitKway,
    PastixSplitKwayProjections
} pastix_split_t;

typedef enum pastix_layout_e {
    PastixRowMajor = 101,
    PastixColMajor = 102
} pastix_layout_t;

typedef enum pastix_trans_e {
    PastixNoTrans   = 111,
    PastixTrans     = 112,
    PastixConjTrans = 113
} pastix_trans_t;

typedef enum pastix_mtxtype_e {
    PastixGeneral   = PastixNoTrans,
    PastixSymmetric = PastixTrans,
    PastixHermitian = PastixConjTrans
} pastix_mtxtype_t;

typedef enum pastix_uplo_e {
    PastixUpper      = 121,
    PastixLower      = 122,
    PastixUpperLower = 123
} pastix_uplo_t;

typedef enum pastix_coefside_e {
    PastixLCoef  = 0,
    PastixUCoef  = 1,
    PastixLUCoef = 2
} pastix_coefside_t;

typedef enum pastix_diag_e {
    PastixNonUnit = 131,
    PastixUnit    = 132
} pastix_diag_t;

typedef enum pastix_side_e {
    PastixLeft  = 141,
    PastixRight = 142
} pastix_side_t;

typedef enum pastix_normtype_e {
    PastixOneNorm       = 171,
    PastixFrobeniusNorm = 174,
    PastixInfNorm       = 175,
    PastixMaxNorm       = 177
} pastix_normtype_t;

typedef enum pastix_dir_e {
    PastixDirForward  = 391,
    PastixDirBackward = 392
} pastix_dir_t;

#endif
