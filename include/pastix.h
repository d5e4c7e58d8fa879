/**
 * @file pastix.h
 *
 * PaStiX C API served by the MI355X engine (libpastix_amd.so).
 *
 * Every declaration below keeps the signature of the reference
 * include/pastix.h:52-172, so code written against PaStiX links against this
 * library unchanged.  Analysis (ordering, symbolic factorization, splitting)
 * runs on the host; pastix_task_numfact / pastix_task_solve /
 * pastix_task_refine run on one MI355X GPU with the factor resident in HBM.
 *
 * Entry points and the reference interface each one replaces:
 *   pastixInitParam          common/api.c:363-538
 *   pastixInit               common/api.c:845
 *   pastixInitWithAffinity   common/api.c:665-820
 *   pastixFinalize           common/api.c:868-952
 *   pastix_task_analyze      blend/pastix_task_analyze.c:52-102
 *   pastix_task_numfact      sopalin/pastix_task_sopalin.c:563-619
 *   pastix_task_solve        sopalin/pastix_task_solve.c:550-577
 *   pastix_task_refine       refinement/pastix_task_refine.c:226-288
 *   pastix_subtask_order     order/pastix_subtask_order.c:97
 *   pastix_subtask_symbfact  symbol/pastix_subtask_symbfact.c:92
 *   pastix_subtask_reordering symbol/pastix_subtask_reordering.c:59
 *   pastix_subtask_blend     blend/pastix_subtask_blend.c:119
 *   pastix_subtask_spm2bcsc  sopalin/pastix_task_sopalin.c:82-166
 *   pastix_subtask_bcsc2ctab sopalin/pastix_task_sopalin.c:196-326
 *   pastix_subtask_sopalin   sopalin/pastix_task_sopalin.c:355-527
 *   pastix_subtask_applyorder sopalin/pastix_task_solve.c:118-170
 *   pastix_subtask_trsm      sopalin/pastix_task_solve.c:219-293
 *   pastix_subtask_diag      sopalin/pastix_task_solve.c:330-371
 *   pastix_subtask_solve     sopalin/pastix_task_solve.c:408-516
 *   pastix_subtask_refine    refinement/pastix_task_refine.c:115-174
 */
#ifndef PASTIX_AMD_PASTIX_H
#define PASTIX_AMD_PASTIX_H

#include <stdint.h>
#include <stddef.h>
#include "spm.h"

typedef uintptr_t PASTIX_Comm;
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD 0
#endif

#include "pastix/api.h"
#include "pastix/datatypes.h"
#include "pastix/order.h"

#ifdef __cplusplus
extern "C" {
#endif

int pastix( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
            pastix_int_t n, pastix_int_t *colptr, pastix_int_t *row,
            void *avals, pastix_int_t *perm, pastix_int_t *invp,
            void *b, pastix_int_t nrhs, pastix_int_t *iparm, double *dparm );

void pastixInitParam( pastix_int_t *iparm, double *dparm );
void pastixInit     ( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
                      pastix_int_t *iparm, double *dparm );
void pastixInitWithAffinity( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
                             pastix_int_t *iparm, double *dparm,
                             const int *bindtab );
void pastixFinalize ( pastix_data_t **pastix_data );

int pastix_task_analyze( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_task_numfact( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_task_solve  ( pastix_data_t *pastix_data, pastix_int_t nrhs,
                         void *b, pastix_int_t ldb );
int pastix_task_refine ( pastix_data_t *pastix_data, pastix_int_t n,
                         pastix_int_t nrhs, void *b, pastix_int_t ldb,
                         void *x, pastix_int_t ldx );

int pastix_subtask_order     ( pastix_data_t *pastix_data, const spmatrix_t *spm,
                               pastix_order_t *myorder );
int pastix_subtask_symbfact  ( pastix_data_t *pastix_data );
int pastix_subtask_reordering( pastix_data_t *pastix_data );
int pastix_subtask_blend     ( pastix_data_t *pastix_data );

int pastix_subtask_spm2bcsc  ( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_subtask_bcsc2ctab ( pastix_data_t *pastix_data );
int pastix_subtask_sopalin   ( pastix_data_t *pastix_data );

int pastix_subtask_applyorder( pastix_data_t *pastix_data, pastix_coeftype_t flttype,
                               pastix_dir_t dir, pastix_int_t m, pastix_int_t n,
                               void *b, pastix_int_t ldb );
int pastix_subtask_trsm( pastix_data_t *pastix_data, pastix_coeftype_t flttype,
                         pastix_side_t side, pastix_uplo_t uplo,
                         pastix_trans_t trans, pastix_diag_t diag,
                         pastix_int_t nrhs, void *b, pastix_int_t ldb );
int pastix_subtask_diag( pastix_data_t *pastix_data, pastix_coeftype_t flttype,
                         pastix_int_t nrhs, void *b, pastix_int_t ldb );
int pastix_subtask_solve( pastix_data_t *pastix_data, pastix_int_t nrhs,
                          void *b, pastix_int_t ldb );
int pastix_subtask_refine( pastix_data_t *pastix_data, pastix_int_t n,
                           pastix_int_t nrhs, const void *b, pastix_int_t ldb,
                           void *x, pastix_int_t ldx );

void pastixSetSchurUnknownList( pastix_data_t *pastix_data, pastix_int_t n,
                                const pastix_int_t *list );
int  pastixGetSchur( const pastix_data_t *pastix_data, void *S, pastix_int_t lds );
void pastixExpand  ( const pastix_data_t *pastix_data, spmatrix_t *spm );
int  pastixGetDiag ( const pastix_data_t *pastix_data, void *D, pastix_int_t incD );

void pastixGetOptions( int argc, char **argv, pastix_int_t *iparm, double *dparm,
                       int *check, spm_driver_t *driver, char **filename );

#ifdef __cplusplus
}
#endif

#endif
