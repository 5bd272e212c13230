/*
 * ORACLE harness (test infrastructure only): linked against object files
 * compiled from the reference's own src/PhysicsData.c and src/Particle.c
 * where they lie, it calls only their pure functions (CrossSection,
 * FUNC_decay, F and the O2sigma* fits) and prints the argon, oxygen and
 * argon/oxygen cross-section tables and the Maxwellian inverse-CDF table as
 * Argon_CrossSectionSET, Oxygen_CrossSectionSET, ArO2_CrossSectionSET
 * (src/PhysicsData.c:1579-3275) and maxwellv (src/Particle.c:176-201) build
 * them.  Those setters write reference globals, so their loops are driven
 * from here with the reference's own arithmetic types.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void CrossSection(int *Cxnum, int *CXsize, float *CXx, float *CXy);
float FUNC_decay(float *m_x, float *m_y, int m_size, float in_x);
float F(float v);
/* the reference's oxygen fits (src/PhysicsData.c:787-1578) */
float O2sigmaE(float), O2sigmaR(float), O2sigmaE1(float), O2sigmaE2(float), O2sigmaE3(float), O2sigmaE4(float);
float O2sigmaSD(float), O2sigmaSS(float), O2sigmaL1(float), O2sigmaL2(float), O2sigmaL3(float), O2sigmaL4(float);
float O2sigmaI(float), O2sigmaLE(float), O2sigmaDA(float), O2sigmaPD(float), O2sigmaDISSIZ(float), O2sigmaEO(float);
float O2sigmaE1D(float), O2sigmaE1S(float), O2sigmaE3P0(float), O2sigmaE5S0(float), O2sigmaE3S0(float);
float O2sigmaIO(float), O2sigmaeD(float), O2sigmaneD(float), O2sigmanegOD(float), O2sigmaDR(float);
float O2sigmaMN(float), O2sigmaMNO(float), O2sigmaCE(float), O2sigmaCEOO(float), O2sigmaCEO(float);
float O2sigmaCEO2O(float), O2sigmaFRAG(float), O2sigmaneE(float), O2sigmapoE(float), O2sigmapooE(float);

/* tabulated data sets, loaded once */
static float sx[17][500], sy[17][500];
static int sn[17];
static float fd(int set, float e) { return FUNC_decay(sx[set], sy[set], sn[set], e); }

/* threshold test of the table builders: below Th_e the entry is 0 */
#define ROW(k, th, expr) do { th_out[k] = (float)(th); c[k] = (e < th_out[k]) ? 0.0f : (float)(expr); } while (0)

/* Oxygen_CrossSectionSET (src/PhysicsData.c:1748-2404) with every reaction
 * switch at 1: one row of 58 cross sections at energy e */
static void o2_row(float e, float *c, float *th_out) {
    const float s1 = 0.977, s2 = 1.627, s3 = 1.96, flag = 1.0f;
    ROW(0, 0.0, flag * O2sigmaE(e));
    ROW(1, 0.07, flag * O2sigmaR(e));
    ROW(2, 0.19, flag * (O2sigmaE1(e) + O2sigmaE2(e) + O2sigmaE3(e) + O2sigmaE4(e)));
    ROW(3, 0.977, flag * O2sigmaSD(e));
    ROW(4, 1.627, flag * O2sigmaSS(e));
    ROW(5, 4.05, flag * O2sigmaL1(e));
    ROW(6, 4.2, flag * O2sigmaDA(e));
    ROW(7, 6.0, flag * O2sigmaL2(e));
    ROW(8, 8.4, flag * O2sigmaL3(e));
    ROW(9, 10.0, flag * O2sigmaL4(e));
    ROW(10, 12.06, flag * O2sigmaI(e));
    ROW(11, 14.7, flag * O2sigmaLE(e));
    ROW(12, 15.0, flag * O2sigmaPD(e));
    ROW(13, 18.73, flag * O2sigmaDISSIZ(e));
    ROW(14, 12.06 - s1, flag * O2sigmaI(e + s1));
    ROW(15, 4.2 - s1, flag * 3.5 * O2sigmaDA(e + s1));
    ROW(16, 0.0, flag * fd(10, e));
    ROW(17, 4.05 - s1, flag * O2sigmaL1(e + s1));
    ROW(18, 6.0 - s1, flag * O2sigmaL2(e + s1));
    ROW(19, 8.4 - s1, flag * O2sigmaL3(e + s1));
    ROW(20, 10.0 - s1, flag * O2sigmaL4(e + s1));
    ROW(21, 18.73 - s1, flag * O2sigmaDISSIZ(e + s1));
    ROW(22, 12.06 - s2, flag * O2sigmaI(e + s2));
    ROW(23, 4.2 - s2, flag * O2sigmaDA(e + s2));
    ROW(24, 0.0, flag * fd(11, e));
    ROW(25, 4.05 - s2, flag * O2sigmaL1(e + s2));
    ROW(26, 6.0 - s2, flag * O2sigmaL2(e + s2));
    ROW(27, 8.4 - s2, flag * O2sigmaL3(e + s2));
    ROW(28, 10.0 - s2, flag * O2sigmaL4(e + s2));
    ROW(29, 18.73 - s2, flag * O2sigmaDISSIZ(e + s2));
    ROW(30, 1.465, flag * O2sigmaeD(e));
    ROW(31, 0.00, flag * O2sigmaDR(e));
    ROW(32, 0.0, flag * O2sigmaEO(e));
    ROW(33, 1.96, flag * O2sigmaE1D(e));
    ROW(34, 4.18, flag * O2sigmaE1S(e));
    ROW(35, 9.14, flag * O2sigmaE5S0(e));
    ROW(36, 9.51, flag * O2sigmaE3S0(e));
    ROW(37, 13.62, flag * O2sigmaIO(e));
    ROW(38, 15.65, flag * O2sigmaE3P0(e));
    ROW(39, 13.62 - s3, flag * O2sigmaIO(e + s3));
    ROW(40, 0.0, flag * fd(12, e));
    ROW(41, 0.0, flag * O2sigmaneE(e));
    ROW(42, 1.465, flag * O2sigmaneD(e));
    ROW(43, 0.0, flag * O2sigmanegOD(e));
    ROW(44, 0.0, flag * O2sigmaMN(e));
    ROW(45, 0.0, flag * O2sigmaMNO(e));
    ROW(46, 0.0, flag * fd(13, e));
    ROW(47, 1.42, flag * O2sigmaCEO2O(e));
    ROW(48, 0.0, flag * O2sigmaCE(e));
    ROW(49, 0.0, flag * O2sigmapoE(e));
    ROW(50, 13.8, flag * O2sigmaFRAG(e));
    ROW(51, 0.0, flag * O2sigmaCE(e));
    ROW(52, 0.0, flag * O2sigmaCE(e));
    ROW(53, 0.0, flag * O2sigmaCEOO(e));
    ROW(54, 0.0, flag * O2sigmapooE(e));
    ROW(55, 0.0, flag * O2sigmaCEO(e));
    ROW(56, 0.0, flag * O2sigmaCEOO(e));
    ROW(57, 0.0, flag * O2sigmaCEOO(e));
}

/* ArO2_CrossSectionSET (src/PhysicsData.c:2446-3254): argon rows 0-4, the
 * oxygen rows 0-52 as 5-57, O2+ + Ar as 58-59, the oxygen rows 53-57 as
 * 60-64, then Ar+ + Ar and Ar+ + O2 as 65-67 */
static void aro2_row(float e, float *c, float *th_out) {
    const float flag = 1.0f;
    float o[58], oth[58];
    ROW(0, 0.0, flag * fd(1, e));
    ROW(1, 11.55, flag * fd(2, e) - flag * fd(3, e));
    ROW(2, 11.548, flag * fd(3, e));
    ROW(3, 15.76, flag * fd(4, e));
    ROW(4, 4.22, flag * fd(5, e));
    o2_row(e, o, oth);
    for (int k = 0; k < 53; ++k) { c[5 + k] = o[k]; th_out[5 + k] = oth[k]; }
    ROW(58, 0.0, flag * fd(14, e));
    ROW(59, 0.0, flag * fd(15, e));
    for (int k = 53; k < 58; ++k) { c[7 + k] = o[k]; th_out[7 + k] = oth[k]; }
    ROW(65, 0.0, flag * fd(7, e));
    ROW(66, 0.0, flag * fd(9, e));
    ROW(67, 0.0, flag * fd(16, e));
}

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
    /* oxygen and argon/oxygen tables on the same grid */
    for (int set = 0; set <= 16; ++set) { int num = set; sn[set] = 0; CrossSection(&num, &sn[set], sx[set], sy[set]); }
    {
        float c[68], th[68];
        printf("O2TABLE %d 58\n", N);
        for (i = 0; i < N; ++i) {
            o2_row(xee[i], c, th);
            for (int k = 0; k < 58; ++k) printf("%.9g%c", c[k], k == 57 ? '\n' : ' ');
        }
        printf("O2TH\n");
        for (int k = 0; k < 58; ++k) printf("%.9g%c", th[k], k == 57 ? '\n' : ' ');
        printf("AO2TABLE %d 68\n", N);
        for (i = 0; i < N; ++i) {
            aro2_row(xee[i], c, th);
            for (int k = 0; k < 68; ++k) printf("%.9g%c", c[k], k == 67 ? '\n' : ' ');
        }
        printf("AO2TH\n");
        for (int k = 0; k < 68; ++k) printf("%.9g%c", th[k], k == 67 ? '\n' : ' ');
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
