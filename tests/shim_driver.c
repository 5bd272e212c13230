/* A C caller of libpicmcc's reference-named shims, as the reference's own
 * driver would bind them: open the default case from a deck file, run the
 * setup calls in the reference's order (src/main.c:45-56, which ends in
 * main_cuda; src/cuda_main.cu:11-60 for the Set_* sequence), bind the stage
 * table as start_cuda does (inc/main.h:34-41, src/cuda_start.cu:2-34), then
 * either step through the table by hand or call main_cuda.  The state after
 * the run is written to a binary file for the test to compare with the
 * pic_* path.
 *
 * usage: shim_driver <deck.json> <precision 0|1> <staged|main> <nsteps> <out.bin>
 * ("main" ignores nsteps: main_cuda runs the deck's RFCycles cycles)
 *
 * Output: per species, int64 n, then x, y, vx, vy, vz (n reals each), id
 * (n uint32), cell_start (G + 1 int32); then the node arrays "phi" and
 * "mccr" as int64 byte count + bytes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "picmcc.h"

static void (*EFIELD)(void), (*MOVE)(void), (*SORT_BOUNDARY)(void), (*MCC)(void), (*DEPOSIT)(void), (*DIAG)(void);

static void die(const char *what) {
    fprintf(stderr, "shim_driver: %s: %s\n", what, pic_last_error());
    exit(2);
}

static void put_array(FILE *f, pic_ctx *c, const char *name) {
    size_t n = 0;
    int es = 0, fl = 0;
    if (pic_array_info(c, name, &n, &es, &fl)) die(name);
    long long bytes = (long long)n * es;
    void *buf = malloc(bytes ? bytes : 1);
    if (pic_get_array(c, name, buf, (size_t)bytes)) die(name);
    fwrite(&bytes, sizeof bytes, 1, f);
    fwrite(buf, 1, (size_t)bytes, f);
    free(buf);
}

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s deck.json precision staged|main nsteps out.bin\n", argv[0]);
        return 2;
    }
    const int prec = atoi(argv[2]);
    const int staged = strcmp(argv[3], "staged") == 0;
    const int nsteps = atoi(argv[4]);
    if (pic_shim_open(argv[1], prec)) die("pic_shim_open");

    /* the reference's setup sequence; the case is already built, so these
     * only check that it is open */
    info_Device();
    start_cuda();
    Set_Device_Parameter();
    Set_MatrixPCG_cuda();
    Set_Particle_cuda();
    Set_NullCollisionTime_cuda();
    Set_SortBoundary_cuda();
    PCG_SOLVER_Laplace();

    pic_ctx *c = pic_shim_ctx();
    pic_info inf;
    if (!c || pic_info_get(c, &inf)) die("pic_info_get");
    EFIELD = Efield_cuda_Basic;
    MOVE = Move_cuda;
    SORT_BOUNDARY = SortBounndary_cuda;
    MCC = inf.nsp == 2 ? MCC_Ar_cuda : inf.nsp == 4 ? MCC_O2_cuda : MCC_ArO2_cuda;
    DEPOSIT = Deposit_cuda;
    DIAG = Diagnostic;

    if (staged) {
        for (int i = 0; i < nsteps; ++i) {
            if (pic_advance_time(c)) die("pic_advance_time");   /* t += dt; tstep++ */
            PCG_SOLVER();
            EFIELD();
            MOVE();
            SORT_BOUNDARY();
            MCC();
            DEPOSIT();
            DIAG();
        }
    } else {
        main_cuda();
    }
    if (pic_sync(c)) die("pic_sync");
    if (pic_info_get(c, &inf)) die("pic_info_get");

    FILE *f = fopen(argv[5], "wb");
    if (!f) {
        perror(argv[5]);
        return 2;
    }
    const int G = inf.ngx * inf.ngy;
    const size_t rs = prec == PIC_FP64_PLAIN ? sizeof(double) : sizeof(float);
    for (int s = 0; s < inf.nsp; ++s) {
        const long long n = inf.np[s];
        void *a[5];
        for (int q = 0; q < 5; ++q) a[q] = malloc(n ? n * rs : 1);
        uint32_t *id = malloc(n ? n * sizeof(uint32_t) : 1);
        int32_t *cs = malloc((size_t)(G + 1) * sizeof(int32_t));
        if (pic_get_particles(c, s, a[0], a[1], a[2], a[3], a[4], id, cs)) die("pic_get_particles");
        fwrite(&n, sizeof n, 1, f);
        for (int q = 0; q < 5; ++q) fwrite(a[q], rs, (size_t)n, f);
        fwrite(id, sizeof(uint32_t), (size_t)n, f);
        fwrite(cs, sizeof(int32_t), (size_t)G + 1, f);
        for (int q = 0; q < 5; ++q) free(a[q]);
        free(id);
        free(cs);
    }
    put_array(f, c, "phi");
    put_array(f, c, "mccr");
    fclose(f);
    printf("shim_driver: %s run done, tstep %lld\n", staged ? "staged" : "main_cuda", inf.tstep);
    pic_shim_close();
    return 0;
}
