/*
 * mask_ref_main.c -- harness that dumps the reference's node-code arrays.
 *
 * Test infrastructure only: linked with the reference's own solid_init.c,
 * slip_ridges_init.c, slip_posts_init.c, slip_spanwise_init.c and globals.c
 * (built in place by oracle/build_mask_ref.sh), it allocates V.s, V.sls and V.sus
 * with the sizes of main (dns-q-mgrd-all-patterns-xyc.c:145-151, 473-479), builds
 * the one-rank Cartesian communicator as main does (:84-88), calls solid_init and
 * the pattern routine selected in definitions.h (:964-973), and writes
 *   int32 xdc ydc zdc xdf ydf zdf, then V.s [xdc][ydc][zdc], V.sls and V.sus
 *   [xdf][ydf][zdf]
 * to the file named on the command line.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "definitions.h"
#include "fun_defs.h"

int main(int argc, char **argv) {
    int dim[2] = {XDIM, YDIM}, period[2] = {1, 1}, rank, size;
    int xdc, ydc, zdc, xdf, ydf, zdf, hdr[6], GR = gratio;  /* main:134 */
    POINTER V;
    PDATA pd;
    FILE *f;
    if (argc != 2) {
        fprintf(stderr, "usage: %s <out>\n", argv[0]);
        return 2;
    }
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (size != XDIM * YDIM) MPI_Abort(MPI_COMM_WORLD, 1);
    MPI_Cart_create(MPI_COMM_WORLD, 2, dim, period, 1, &cart_grid);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    memset(&pd, 0, sizeof pd);
    pd.myrank = rank;
    pd.numproc = size;
    xdc = 1 + (xdv / XDIM);
    ydc = 1 + (ydv / YDIM);
    zdc = zdv;
    xdf = GR * (xdc - 1) + 1;
    ydf = GR * (ydc - 1) + 1;
    zdf = zds + 1;
    memset(&V, 0, sizeof V);
    V.s = (int *)calloc((size_t)xdc * ydc * zdc, sizeof(int));
    V.sls = (int *)calloc((size_t)xdf * ydf * zdf, sizeof(int));
    V.sus = (int *)calloc((size_t)xdf * ydf * zdf, sizeof(int));
    if (!V.s || !V.sls || !V.sus) MPI_Abort(MPI_COMM_WORLD, 1);
    V = solid_init(xdc, ydc, zdc, xdf, ydf, zdf, V);
#ifdef RSLIP
    V = slip_ridges_init(xdf, ydf, zdf, V, pd);
#endif
#ifdef PSLIP
    V = slip_posts_init(xdf, ydf, zdf, V, pd);
#endif
#ifdef SSLIP
    V = slip_spanwise_init(xdf, ydf, zdf, V, pd);
#endif
    hdr[0] = xdc; hdr[1] = ydc; hdr[2] = zdc; hdr[3] = xdf; hdr[4] = ydf; hdr[5] = zdf;
    f = fopen(argv[1], "wb");
    if (!f || fwrite(hdr, sizeof(int), 6, f) != 6 ||
        fwrite(V.s, sizeof(int), (size_t)xdc * ydc * zdc, f) != (size_t)xdc * ydc * zdc ||
        fwrite(V.sls, sizeof(int), (size_t)xdf * ydf * zdf, f) != (size_t)xdf * ydf * zdf ||
        fwrite(V.sus, sizeof(int), (size_t)xdf * ydf * zdf, f) != (size_t)xdf * ydf * zdf) {
        fprintf(stderr, "write failed\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    fclose(f);
    MPI_Finalize();
    return 0;
}
