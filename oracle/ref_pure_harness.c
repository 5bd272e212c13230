/*
 * ORACLE harness (test infrastructure only): linked against object files
 * compiled from the reference's own src/PhysicsData.c and src/Particle.c
 * where they lie, it calls only their pure functions (CrossSection,
 * FUNC_decay, F) and prints the argon cross-section table and the
 * Maxwellian inverse-CDF table exactly as Argon_CrossSectionSET
 * (src/PhysicsData.c:1579-1718) and maxwellv (src/Particle.c:176-201) build
 * them.  Those two setters write reference globals, so their loops are
 * driven from here with the reference's own arithmetic types.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void CrossSection(int *Cxnum, int *CXsize, float *CXx, float *CXy);
float FUNC_decay(float *m_x, float *m_y, int m_size, float in_x);
float F(float v);

static float bx[500], by[500], bx2[500], by2[500];

int main(int argc, char **argv) {
    int i, n, s2;
    /* raw data sets */
    for (int set = 0; set <= 16; ++set) {
        int num = set, size = 0;
        CrossSection(&num, &size, bx, by);
        printf("SET %d %d\n", set, size);
        for (i = 0; i < size; ++i) printf("%.9g %.9g\n", bx[i], by[i]);
    }
    /* argon table: src/PhysicsData.c:1586-1713 with all flags = 1 */
    float LOGX_MIN = -4.0f, LOGX_MAX = 4.0f, dLOGX = 0.1f, idLOGX = 1 / dLOGX;
    int N = (int)((LOGX_MAX - LOGX_MIN) * idLOGX) + 1;
    float xe[200], xee[200];
    xe[0] = idLOGX * LOGX_MIN;
    for (i = 1; i < N; ++i) xe[i] = xe[i - 1] + 1;
    for (i = 0; i < N; ++i) { xe[i] = dLOGX * xe[i]; xee[i] = pow(10, xe[i]); }
    printf("GRID %d %.9g\n", N, idLOGX);
    for (i = 0; i < N; ++i) printf("%.9g %.9g\n", xe[i], xee[i]);
    const int dn[7] = {1, 2, 3, 4, 5, 7, 9};
    const float th[7] = {0.0f, 11.55f, 11.548f, 15.76f, 4.22f, 0.0f, 0.0f};
    const float flag = 1.0f;
    for (int r = 0; r < 7; ++r) {
        int d = dn[r], size = 0, size2 = 0, d2 = 3;
        CrossSection(&d, &size, bx, by);
        if (r == 1) CrossSection(&d2, &size2, bx2, by2);
        printf("AR %d\n", r);
        for (i = 0; i < N; ++i) {
            float v;
            if (xee[i] < th[r]) v = 0.0f;
            else if (r == 1) v = flag * FUNC_decay(bx, by, size, xee[i]) - flag * FUNC_decay(bx2, by2, size2, xee[i]);
            else v = flag * FUNC_decay(bx, by, size, xee[i]);
            printf("%.9g\n", v);
        }
    }
    /* Maxwellian inverse CDF: src/Particle.c:184-199 */
    int nvel = (int)(1. / (1 - F((float)3.3)));
    float dv = sqrt(3.14159265358979323846) / (4.0 * nvel);
    printf("VSAVE %d\n", nvel);
    i = 0;
    for (n = 0; n < nvel; ++n) {
        float rr = (1.0 * n) / nvel;
        while (F(i * dv) < rr) i++;
        printf("%.9g\n", (float)(sqrt(2.0) * i * dv));
    }
    (void)s2; (void)argc; (void)argv;
    return 0;
}
