/**
 * @file pastix/order.h
 *
 * Ordering structure and routines of the PaStiX C API.
 *
 * pastix_order_t has the field order of the reference
 * include/pastix/order.h:40-51 so a caller-built ordering (Personal ordering,
 * pastixOrderGrid) can be passed to pastix_subtask_order.
 */
#ifndef PASTIX_AMD_ORDER_H
#define PASTIX_AMD_ORDER_H

#include "pastix/datatypes.h"

typedef struct pastix_order_s {
    pastix_int_t  baseval;   /* base value used for numbering                 */
    pastix_int_t  vertnbr;   /* number of vertices                            */
    pastix_int_t  cblknbr;   /* number of column blocks (supernodes)          */
    pastix_int_t *permtab;   /* permutation, size vertnbr                     */
    pastix_int_t *peritab;   /* inverse permutation, size vertnbr             */
    pastix_int_t *rangtab;   /* supernode ranges, size cblknbr+1              */
    pastix_int_t *treetab;   /* supernode tree (father index, -1 root)        */
    int8_t       *selevtx;   /* low-rank clustering flags (unused, NULL)      */
    pastix_int_t  sndenbr;   /* number of supernodes before splitting         */
    pastix_int_t *sndetab;   /* original supernode ranges, size sndenbr+1     */
} pastix_order_t;

#ifdef __cplusplus
extern "C" {
#endif

/* reference order/order.c */
int  pastixOrderInit ( pastix_order_t * const ordeptr, pastix_int_t baseval,
                       pastix_int_t vertnbr, pastix_int_t cblknbr,
                       pastix_int_t * const permtab, pastix_int_t * const peritab,
                       pastix_int_t * const rangtab, pastix_int_t * const treetab );
int  pastixOrderAlloc( pastix_order_t * const ordeptr, pastix_int_t vertnbr,
                       pastix_int_t cblknbr );
void pastixOrderExit ( pastix_order_t * const ordeptr );
void pastixOrderBase ( pastix_order_t * const ordeptr, pastix_int_t baseval );
int  pastixOrderCheck( const pastix_order_t * const ordeptr );

/* reference order/order_io.c:132-181,294-330: the file named by the
 * PASTIX_FILE_ORDER environment variable (default "ordername" to load,
 * "ordergen" in the instance's output directory to save) */
int  pastixOrderLoad( const pastix_data_t *pastix_data, pastix_order_t *ordeptr );
int  pastixOrderSave( pastix_data_t *pastix_data, const pastix_order_t *ordeptr );

/* reference order/order_grids.c:275-383 */
int  pastixOrderGrid ( pastix_order_t **myorder, pastix_int_t nx,
                       pastix_int_t ny, pastix_int_t nz );

/* reference order/order_apply_level_order.c:94 */
int  pastixOrderApplyLevelOrder( pastix_order_t *ordeptr,
                                 pastix_int_t level_tasks2d,
                                 pastix_int_t width_tasks2d );

#ifdef __cplusplus
}
#endif

#endif
