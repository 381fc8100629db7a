/*
 * rrlbm_oracle.c -- CPU restatement of the reference RR-LBM routines.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may load it; the product path (the HIP library and the
 * rrlbm Python package) never calls it and has no CPU fallback.
 *
 * Every routine restates one reference routine (file:line relative to the
 * reference repository root) in plain C with FP64 arithmetic, keeping the
 * reference's loop structure (per-direction Hermite tables, in-place
 * collision, push streaming, separate bounce-back / macrovar / RRBC passes),
 * so that the optimised GPU formulation (moment space, pull streaming, fused)
 * is checked against an independent formulation.
 *
 * Arrays use the channel solver's layout: populations SoA f[ip*N + idx],
 * idx = (iz*LY + iy)*LX + ix (RR27V33_code_DNS_LES/collision.cu:11,73).
 * The Taylor-Green restatement keeps the Fortran layout f(ip,i,j,k)
 * (ip fastest, column-major) and ordering.
 *
 * Pinning (each function's check lives in tests/test_oracle_*.py):
 *   lattice / Hermite tables  -> closed-form moments and orthogonality
 *   orc_feq                   -> exact 0th-3rd lattice moments
 *   orc_collision_rr          -> mass conservation, momentum change = F,
 *                                projection idempotence, tau=1 / u=0 cases
 *   orc_streaming, orc_bb     -> permutation / delta tests, closed-box mass
 *   orc_macrovar(_bn)         -> closed form
 *   orc_rrbc                  -> outgoing populations untouched; equilibrium in
 *                                -> equilibrium out
 *   orc_les_wale              -> pure shear gives tau_t = 0; long-double brute
 *                                force on random linear fields
 *   orc_tg_ref_*              -> committed Fortran output
 *                                Taylor_Green/RRBMa0.01L100.txt (20 000 steps)
 *   orc_step_laminar          -> committed FP64 fields Laminar/output/uy*.dat
 *                                (5 000 and 10 000 steps, 1e-14 relative)
 *   sixth-order RR terms       -> tables: orthogonality of all 27 tensor
 *   (orc_set_rr_order(6))        Hermite polynomials; feq = w rho prod_a
 *                                (1 + 3 u_a c_a + 9/2 u_a^2 c_a^2 - 3/2 u_a^2);
 *                                the non-equilibrium part against its
 *                                generating-function form (second derivatives
 *                                of that product), mass and momentum balance
 *   orc_fdbc                  -> closed forms: uniform flow gives the
 *                                equilibrium at 7/15 u with the known-sum
 *                                density; linear and parabolic profiles are
 *                                extrapolated exactly; the reconstructed
 *                                second moment is the Chapman-Enskog stress
 *                                -2/3 rho tau S; idempotence
 *   orc_collision_bgk,        -> mass conservation, momentum change = F,
 *   orc_macrovar_bgk             the rest state as a fixed point, the
 *                                second-moment relaxation rate 1 - 1/tau
 *                                ("parity unpinned" beyond that: the BGK
 *                                solver's repository has no committed fields)
 *   orc_forcing_gp, orc_init_arrays -> closed-form spot values of the cited
 *                                formulas ("parity unpinned" beyond that: no
 *                                committed reference output exists)
 */
#include <omp.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_CS 0.57735026918962576451 /* 1/sqrt(3) */

/* ------------------------------------------------------------------ */
/* Lattices                                                            */
/* ------------------------------------------------------------------ */
enum { ORC_D3Q27_CUDA = 0, ORC_D3Q27_FORTRAN = 1, ORC_D3Q19 = 2 };

typedef struct {
    int q;
    int cx[27], cy[27], cz[27], opp[27];
    double w[27];
    /* scaled Hermite tables, RR27V33_code_DNS_LES/initial.cu:92-110 */
    double H100[27], H010[27], H001[27];
    double H110[27], H101[27], H011[27], H200[27], H020[27], H002[27];
    double H210[27], H201[27], H021[27], H120[27], H102[27], H012[27], H111[27];
    /* fourth to sixth order (the commented tables of initial.cu:111-130) */
    double H220[27], H202[27], H022[27], H211[27], H121[27], H112[27];
    double H221[27], H212[27], H122[27], H222[27];
} orc_lat;

static void lat_build(int kind, orc_lat* L) {
    /* RR27V33_code_DNS_LES/initial.cu:67-77 */
    static const int cx27c[27] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static const int cy27c[27] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    static const int cz27c[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    /* Taylor_Green/D3Q27RR.f90:69-71 (first 19 entries: D3Q19RR.f90:75-77) */
    static const int cx27f[27] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1};
    static const int cy27f[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    static const int cz27f[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};
    const int* X = kind == ORC_D3Q27_CUDA ? cx27c : cx27f;
    const int* Y = kind == ORC_D3Q27_CUDA ? cy27c : cy27f;
    const int* Z = kind == ORC_D3Q27_CUDA ? cz27c : cz27f;
    L->q = kind == ORC_D3Q19 ? 19 : 27;
    double RT = 1.0 / 3.0, cs = sqrt(RT), cs3 = RT * cs;
    for (int i = 0; i < L->q; ++i) {
        int x = X[i], y = Y[i], z = Z[i];
        L->cx[i] = x; L->cy[i] = y; L->cz[i] = z;
        int nz = (x != 0) + (y != 0) + (z != 0);
        if (L->q == 27)
            L->w[i] = nz == 0 ? 8.0 / 27.0 : nz == 1 ? 2.0 / 27.0 : nz == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
        else
            L->w[i] = nz == 0 ? 1.0 / 3.0 : nz == 1 ? 1.0 / 18.0 : 1.0 / 36.0;
        L->H100[i] = x / cs; L->H010[i] = y / cs; L->H001[i] = z / cs;
        L->H110[i] = x * y / RT; L->H101[i] = x * z / RT; L->H011[i] = y * z / RT;
        L->H200[i] = x * x / RT - 1; L->H020[i] = y * y / RT - 1; L->H002[i] = z * z / RT - 1;
        L->H210[i] = x * x * y / cs3 - y / cs;
        L->H201[i] = x * x * z / cs3 - z / cs;
        L->H021[i] = y * y * z / cs3 - z / cs;
        L->H120[i] = x * y * y / cs3 - x / cs;
        L->H102[i] = x * z * z / cs3 - x / cs;
        L->H012[i] = y * z * z / cs3 - y / cs;
        L->H111[i] = x * y * z / cs3;
        /* initial.cu:111-130 (commented out there): products of the 1D
         * polynomials 1, c/cs, c^2/cs^2 - 1.  The reference's H222 line has
         * + c_y^2 c_z^2 / cs^4 where the product (and orthogonality to H000,
         * hence mass conservation) needs -; this restatement uses -. */
        double X2 = x * x / RT, Y2 = y * y / RT, Z2 = z * z / RT;
        double X1 = x / cs, Y1 = y / cs, Z1 = z / cs;
        L->H220[i] = X2 * Y2 - X2 - Y2 + 1;
        L->H202[i] = X2 * Z2 - X2 - Z2 + 1;
        L->H022[i] = Y2 * Z2 - Y2 - Z2 + 1;
        L->H211[i] = X2 * Y1 * Z1 - Y1 * Z1;
        L->H121[i] = X1 * Y2 * Z1 - X1 * Z1;
        L->H112[i] = X1 * Y1 * Z2 - X1 * Y1;
        L->H221[i] = X2 * Y2 * Z1 - X2 * Z1 - Y2 * Z1 + Z1;
        L->H212[i] = X2 * Y1 * Z2 - X2 * Y1 - Y1 * Z2 + Y1;
        L->H122[i] = X1 * Y2 * Z2 - X1 * Y2 - X1 * Z2 + X1;
        L->H222[i] = X2 * Y2 * Z2 - X2 * Y2 - X2 * Z2 - Y2 * Z2 + X2 + Y2 + Z2 - 1;
    }
    for (int i = 0; i < L->q; ++i)
        for (int j = 0; j < L->q; ++j)
            if (L->cx[j] == -L->cx[i] && L->cy[j] == -L->cy[i] && L->cz[j] == -L->cz[i]) L->opp[i] = j;
}

void orc_lattice(int kind, int* cx, int* cy, int* cz, double* w, int* opp) {
    orc_lat L;
    lat_build(kind, &L);
    for (int i = 0; i < L.q; ++i) {
        cx[i] = L.cx[i]; cy[i] = L.cy[i]; cz[i] = L.cz[i]; w[i] = L.w[i]; opp[i] = L.opp[i];
    }
}

/* Hermite tables for direction i, in the order
 * H100 H010 H001 H110 H101 H011 H200 H020 H002 H210 H201 H021 H120 H102 H012 H111 */
void orc_hermite(int kind, double* out /* 16*q */) {
    orc_lat L;
    lat_build(kind, &L);
    for (int i = 0; i < L.q; ++i) {
        double* o = out + 16 * i;
        o[0] = L.H100[i]; o[1] = L.H010[i]; o[2] = L.H001[i];
        o[3] = L.H110[i]; o[4] = L.H101[i]; o[5] = L.H011[i];
        o[6] = L.H200[i]; o[7] = L.H020[i]; o[8] = L.H002[i];
        o[9] = L.H210[i]; o[10] = L.H201[i]; o[11] = L.H021[i];
        o[12] = L.H120[i]; o[13] = L.H102[i]; o[14] = L.H012[i]; o[15] = L.H111[i];
    }
}

/* Tables 4th-6th order for direction i, in the order
 * H220 H202 H022 H211 H121 H112 H221 H212 H122 H222 */
void orc_hermite_high(int kind, double* out /* 10*q */) {
    orc_lat L;
    lat_build(kind, &L);
    for (int i = 0; i < L.q; ++i) {
        double* o = out + 10 * i;
        o[0] = L.H220[i]; o[1] = L.H202[i]; o[2] = L.H022[i];
        o[3] = L.H211[i]; o[4] = L.H121[i]; o[5] = L.H112[i];
        o[6] = L.H221[i]; o[7] = L.H212[i]; o[8] = L.H122[i]; o[9] = L.H222[i];
    }
}

/* ------------------------------------------------------------------ */
/* Equilibrium and the RR operator for one cell                        */
/* ------------------------------------------------------------------ */
typedef struct {
    double a110, a101, a011, a200, a020, a002, a210, a201, a120, a021, a012, a102, a111;
    double a220, a202, a022, a211, a121, a112, a221, a212, a122, a222;
} orc_a;

/* Order of the RR expansion: 3 (the reference's active code) or 6 (the
 * fourth- to sixth-order terms the reference carries commented out,
 * collision.cu:48-49,62-67,95-104,112-116 and BC.cu:37-38,50-54,79-88,97-101;
 * D3Q27 only). */
static int g_rr_order = 3;
/* Threads of the OpenMP loops (CPU timing of the oracle). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

void orc_set_rr_order(int order) { g_rr_order = order == 6 ? 6 : 3; }
int orc_get_rr_order(void) { return g_rr_order; }
static int high_order(const orc_lat* L) { return g_rr_order == 6 && L->q == 27; }

/* Third-order Hermite sum for direction ip.
 * D3Q27: collision.cu:60-61.  D3Q19: Malaspinas pairs, D3Q19RR.f90:131-136,230-235. */
static double third_sum(const orc_lat* L, int ip, const orc_a* a) {
    if (L->q == 27)
        return L->H210[ip] * a->a210 + L->H201[ip] * a->a201 + L->H021[ip] * a->a021 + L->H120[ip] * a->a120 +
               L->H012[ip] * a->a012 + L->H102[ip] * a->a102 + 2.0 * L->H111[ip] * a->a111;
    double Hp1 = L->H210[ip] + L->H012[ip], Hm1 = L->H210[ip] - L->H012[ip];
    double Hp2 = L->H102[ip] + L->H120[ip], Hm2 = L->H102[ip] - L->H120[ip];
    double Hp3 = L->H021[ip] + L->H201[ip], Hm3 = L->H021[ip] - L->H201[ip];
    return Hp1 * (a->a210 + a->a012) + Hp2 * (a->a102 + a->a120) + Hp3 * (a->a021 + a->a201) +
           Hm1 * (a->a210 - a->a012) / 3.0 + Hm2 * (a->a102 - a->a120) / 3.0 + Hm3 * (a->a021 - a->a201) / 3.0;
}

static double second_sum(const orc_lat* L, int ip, const orc_a* a) {
    return 2.0 * (L->H110[ip] * a->a110 + L->H101[ip] * a->a101 + L->H011[ip] * a->a011) +
           L->H200[ip] * a->a200 + L->H020[ip] * a->a020 + L->H002[ip] * a->a002;
}

static void eq_coeffs(double uu, double vv, double ww, orc_a* a) {
    /* collision.cu:44-47 */
    a->a110 = uu * vv; a->a101 = uu * ww; a->a011 = vv * ww;
    a->a200 = uu * uu; a->a020 = vv * vv; a->a002 = ww * ww;
    a->a210 = a->a200 * vv; a->a201 = a->a200 * ww; a->a120 = a->a020 * uu; a->a021 = a->a020 * ww;
    a->a012 = a->a002 * vv; a->a102 = a->a002 * uu; a->a111 = a->a110 * ww;
    /* collision.cu:48-49 (commented there) */
    a->a220 = a->a210 * vv; a->a202 = a->a201 * ww; a->a022 = a->a021 * ww;
    a->a211 = a->a210 * ww; a->a121 = a->a120 * ww; a->a112 = a->a102 * vv;
    a->a221 = a->a220 * ww; a->a212 = a->a202 * vv; a->a122 = a->a022 * uu; a->a222 = a->a221 * ww;
}

/* four = H220 a220 + H202 a202 + H022 a022 + 2 (H211 a211 + H121 a121 + H112 a112),
 * enters as four/4 (collision.cu:62-63,66) */
static double fourth_sum(const orc_lat* L, int ip, const orc_a* a) {
    return L->H220[ip] * a->a220 + L->H202[ip] * a->a202 + L->H022[ip] * a->a022 +
           2.0 * (L->H211[ip] * a->a211 + L->H121[ip] * a->a121 + L->H112[ip] * a->a112);
}

/* fifsix = (H221 a221 + H212 a212 + H122 a122)/4 + H222 a222/8 (collision.cu:64) */
static double fifsix_sum(const orc_lat* L, int ip, const orc_a* a) {
    return (L->H221[ip] * a->a221 + L->H212[ip] * a->a212 + L->H122[ip] * a->a122) / 4.0 +
           L->H222[ip] * a->a222 / 8.0;
}

/* feq_i = w_i rho (1 + H1.u~ + (sec + thir)/2)   (collision.cu:57-69) */
static void feq_cell(const orc_lat* L, double rho, double u, double v, double w, double* feq) {
    double uu = u / ORC_CS, vv = v / ORC_CS, ww = w / ORC_CS;
    orc_a a;
    eq_coeffs(uu, vv, ww, &a);
    for (int ip = 0; ip < L->q; ++ip) {
        double fir = L->H100[ip] * uu + L->H010[ip] * vv + L->H001[ip] * ww;
        double hi = high_order(L) ? fourth_sum(L, ip, &a) / 4.0 + fifsix_sum(L, ip, &a) : 0.0;
        feq[ip] = L->w[ip] * rho * (1.0 + fir + 0.5 * (second_sum(L, ip, &a) + third_sum(L, ip, &a)) + hi);
    }
}

void orc_feq(int kind, double rho, double u, double v, double w, double* feq) {
    orc_lat L;
    lat_build(kind, &L);
    feq_cell(&L, rho, u, v, w, feq);
}

/* The RR reconstruction of one cell.
 *   mode 0 (collision, collision.cu:52-118): out_i = feq + (1 - 1/tau_eff) fneq + Fbar feq
 *   mode 1 (RRBC,      BC.cu:41-122):        out_i = feq + fneq - Fbar feq
 * f holds the cell's 27 (or 19) populations; the projection uses f + Fbar feq. */
/* Fbar_i = (c_i.F - u.F)/(2 cs^2), collision.cu:33-39 */
static void fbar_cell(const orc_lat* L, double u, double v, double w, double fx, double fy, double fz, double* Fbar) {
    double G3 = u * fx + v * fy + w * fz;
    for (int ip = 0; ip < L->q; ++ip) {
        double G1 = L->cx[ip] * fx + L->cy[ip] * fy + L->cz[ip] * fz;
        Fbar[ip] = 0.5 * (G1 - G3) / (1.0 / 3.0);
    }
}

/* As rr_cell with the forcing vector Fbar given (the Laminar solver's wall
 * regularisation passes a Fbar computed at another cell). */
static void rr_cell_fbar(const orc_lat* L, const double* f, double rho, double u, double v, double w,
                         const double* Fbar, double tau_eff, int mode, double* out) {
    /* feq: the equilibrium; ffeq: the equilibrium the force term multiplies
     * (forcefeq, collision.cu:66-67: the same at third order, without the
     * fifth/sixth-order terms at sixth order) */
    double feq[27], ffeq[27];
    const int ho = high_order(L);
    double uu = u / ORC_CS, vv = v / ORC_CS, ww = w / ORC_CS;
    orc_a a;
    eq_coeffs(uu, vv, ww, &a);
    double ane110 = 0, ane101 = 0, ane011 = 0, ane200 = 0, ane020 = 0, ane002 = 0;
    for (int ip = 0; ip < L->q; ++ip) {
        double fir = L->H100[ip] * uu + L->H010[ip] * vv + L->H001[ip] * ww;
        double base = 1.0 + fir + 0.5 * (second_sum(L, ip, &a) + third_sum(L, ip, &a));
        if (ho) base += fourth_sum(L, ip, &a) / 4.0;
        ffeq[ip] = L->w[ip] * rho * base;
        feq[ip] = ho ? L->w[ip] * rho * (base + fifsix_sum(L, ip, &a)) : ffeq[ip];
        double g = f[ip] + Fbar[ip] * ffeq[ip];
        ane110 += L->H110[ip] * g; ane101 += L->H101[ip] * g; ane011 += L->H011[ip] * g;
        ane200 += L->H200[ip] * g; ane020 += L->H020[ip] * g; ane002 += L->H002[ip] * g;
    }
    /* collision.cu:81-94 */
    orc_a n;
    n.a110 = ane110 - rho * uu * vv; n.a101 = ane101 - rho * uu * ww; n.a011 = ane011 - rho * vv * ww;
    n.a200 = ane200 - rho * uu * uu; n.a020 = ane020 - rho * vv * vv; n.a002 = ane002 - rho * ww * ww;
    n.a210 = n.a200 * vv + 2. * uu * n.a110;
    n.a201 = n.a200 * ww + 2. * uu * n.a101;
    n.a120 = n.a020 * uu + 2. * vv * n.a110;
    n.a021 = n.a020 * ww + 2. * vv * n.a011;
    n.a102 = n.a002 * uu + 2. * ww * n.a101;
    n.a012 = n.a002 * vv + 2. * ww * n.a011;
    n.a111 = n.a110 * ww + n.a101 * vv + n.a011 * uu;
    if (ho) {
        /* collision.cu:95-104 (commented there).  Fourth order as written; in
         * the fifth/sixth-order lines the reference multiplies by a_200 =
         * rho uu^2 (and a_020, a_110, a_011, a_101) where the recursion of its
         * own lower orders (u products only, as in Malaspinas 2015) has
         * uu^2: this restatement uses the u products, so every order follows
         * one recursion. */
        n.a220 = n.a210 * vv + uu * uu * n.a020 + 2. * uu * vv * n.a110;
        n.a202 = n.a201 * ww + uu * uu * n.a002 + 2. * uu * ww * n.a101;
        n.a022 = n.a021 * ww + vv * vv * n.a002 + 2. * vv * ww * n.a011;
        n.a211 = n.a210 * ww + uu * uu * n.a011 + 2. * uu * vv * n.a101;
        n.a121 = n.a120 * ww + 2. * uu * vv * n.a011 + vv * vv * n.a101;
        n.a112 = n.a111 * ww + uu * vv * n.a002 + uu * ww * n.a011 + vv * ww * n.a101;
        n.a221 = n.a220 * ww + 2. * uu * uu * vv * n.a011 + 2. * vv * vv * uu * n.a101;
        n.a212 = n.a211 * ww + 2. * uu * vv * ww * n.a101 + uu * uu * (vv * n.a002 + ww * n.a011);
        n.a122 = n.a121 * ww + 2. * uu * vv * ww * n.a011 + vv * vv * (uu * n.a002 + ww * n.a101);
        n.a222 = n.a221 * ww + uu * uu * vv * vv * n.a002 + 2. * uu * uu * vv * ww * n.a011 +
                 2. * uu * vv * vv * ww * n.a101;
    }
    for (int ip = 0; ip < L->q; ++ip) {
        double fneq = 0.5 * (second_sum(L, ip, &n) + third_sum(L, ip, &n));
        if (ho) fneq += fourth_sum(L, ip, &n) / 4.0 + fifsix_sum(L, ip, &n); /* collision.cu:112-116 */
        fneq *= L->w[ip];
        if (mode == 0)
            out[ip] = feq[ip] + (1.0 - 1.0 / tau_eff) * fneq + Fbar[ip] * ffeq[ip];
        else
            out[ip] = feq[ip] + fneq - Fbar[ip] * ffeq[ip];
    }
}

static void rr_cell(const orc_lat* L, const double* f, double rho, double u, double v, double w,
                    double fx, double fy, double fz, double tau_eff, int mode, double* out) {
    double Fbar[27];
    fbar_cell(L, u, v, w, fx, fy, fz, Fbar);
    rr_cell_fbar(L, f, rho, u, v, w, Fbar, tau_eff, mode, out);
}

/* One-cell collision on caller arrays (for closed-form tests). */
void orc_rr_cell(int kind, const double* f, double rho, double u, double v, double w, double fx, double fy,
                 double fz, double tau_eff, int mode, double* out) {
    orc_lat L;
    lat_build(kind, &L);
    rr_cell(&L, f, rho, u, v, w, fx, fy, fz, tau_eff, mode, out);
}

/* collision_RR, RR27V33_code_DNS_LES/collision.cu:4-119 (in place). taut may be NULL (USE_LES 0). */
void orc_collision_rr(int kind, int lx, int ly, int lz, double* f, const double* fx, const double* fy,
                      const double* fz, const double* rho, const double* ux, const double* uy, const double* uz,
                      const double* taut, double tau) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; ++idx) {
        double fc[27], out[27];
        for (int ip = 0; ip < L.q; ++ip) fc[ip] = f[ip * N + idx];
        double tau_eff = tau + (taut ? taut[idx] : 0.0); /* collision.cu:27-31 */
        rr_cell(&L, fc, rho[idx], ux[idx], uy[idx], uz[idx], fx[idx], fy[idx], fz[idx], tau_eff, 0, out);
        for (int ip = 0; ip < L.q; ++ip) f[ip * N + idx] = out[ip];
    }
}

/* ------------------------------------------------------------------ */
/* Streaming, bounce-back, moments                                     */
/* ------------------------------------------------------------------ */
static inline long cidx(int ix, int iy, int iz, int lx, int ly) { return ((long)iz * ly + iy) * lx + ix; }

/* streaming, streaming.cu:3-24: push into ftmp, periodic in x, y, z. */
void orc_streaming(int kind, int lx, int ly, int lz, const double* f, double* ftmp) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                for (int ip = 0; ip < L.q; ++ip) {
                    int im = (ix + L.cx[ip] + lx) % lx, jm = (iy + L.cy[ip] + ly) % ly, km = (iz + L.cz[ip] + lz) % lz;
                    ftmp[ip * N + cidx(im, jm, km, lx, ly)] = f[ip * N + idx];
                }
            }
}

/* Face f = 2*axis + side (side 0: low face, 1: high face).  Population ip
 * enters the domain through face f when its source cell lies beyond it. */
static int enters_through(const orc_lat* L, int ip, int face) {
    int axis = face / 2, side = face % 2;
    int c = axis == 0 ? L->cx[ip] : axis == 1 ? L->cy[ip] : L->cz[ip];
    return side == 0 ? c > 0 : c < 0;
}
static int on_face(int face, int ix, int iy, int iz, int lx, int ly, int lz) {
    int axis = face / 2, side = face % 2;
    int p = axis == 0 ? ix : axis == 1 ? iy : iz;
    int n = axis == 0 ? lx : axis == 1 ? ly : lz;
    return side == 0 ? p == 0 : p == n - 1;
}

/* Half-way bounce-back of the wall-incoming populations, streaming.cu:26-55
 * (channel: walls at iy=0 and iy=LY-1), generalised to walls on any axis and a
 * moving lid (Cavity/streaming.f90:91-99,142-152: f_i = f*_opp + 6 w_i c_i.u_lid).
 * Faces are processed in the order x-,x+,y-,y+,z-,z+; a population entering
 * through several faces takes the value of the last one. */
void orc_bb(int kind, int lx, int ly, int lz, const double* f, double* ftmp, int wall_x, int wall_y, int wall_z,
            int lid_face, const double* lid_u) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
    int walled[3] = {wall_x, wall_y, wall_z};
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                for (int face = 0; face < 6; ++face) {
                    if (!walled[face / 2] || !on_face(face, ix, iy, iz, lx, ly, lz)) continue;
                    for (int ip = 0; ip < L.q; ++ip) {
                        if (!enters_through(&L, ip, face)) continue;
                        double v = f[L.opp[ip] * N + idx];
                        if (face == lid_face)
                            v += 6.0 * L.w[ip] * (L.cx[ip] * lid_u[0] + L.cy[ip] * lid_u[1] + L.cz[ip] * lid_u[2]);
                        ftmp[ip * N + idx] = v;
                    }
                }
            }
}

/* macrovar, macrovar.cu:4-31; with wall_only set, macrovar_BN (macrovar.cu:33-62)
 * restricted to cells on a walled face. */
void orc_macrovar(int kind, int lx, int ly, int lz, const double* f, const double* fx, const double* fy,
                  const double* fz, double* rho, double* ux, double* uy, double* uz, int wall_only, int wall_x,
                  int wall_y, int wall_z) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                if (wall_only) {
                    int on = (wall_x && (ix == 0 || ix == lx - 1)) || (wall_y && (iy == 0 || iy == ly - 1)) ||
                             (wall_z && (iz == 0 || iz == lz - 1));
                    if (!on) continue;
                }
                long idx = cidx(ix, iy, iz, lx, ly);
                double r = 0, jx = 0, jy = 0, jz = 0;
                for (int ip = 0; ip < L.q; ++ip) {
                    double v = f[ip * N + idx];
                    r += v; jx += L.cx[ip] * v; jy += L.cy[ip] * v; jz += L.cz[ip] * v;
                }
                rho[idx] = r;
                ux[idx] = (jx + 0.5 * fx[idx]) / r;
                uy[idx] = (jy + 0.5 * fy[idx]) / r;
                uz[idx] = (jz + 0.5 * fz[idx]) / r;
            }
}

/* RRBC, BC.cu:4-126: RR regularisation of the wall-incoming populations,
 * out_i = feq + fneq - Fbar feq, with (rho, u) from the current populations.
 * Generalised to several walled faces as the reference's Cavity solver does
 * it (Cavity/streaming.f90:89-292): face by face in the order x-,x+,y-,y+,
 * z-,z+, each face sets its incoming populations to their bounce-back values
 * (from fpost, the post-collision populations, plus the lid term) and then
 * regularises them with moments recomputed from the cell's current
 * populations.  While an earlier face is treated, the populations entering
 * through a later face still hold the values the periodic streaming brought
 * (from the cell across the domain), as in the Fortran, where that face's
 * bounce-back has not run yet: a cell on several faces first takes those
 * values back from fpost.  For the channel (one face per cell) this is
 * exactly the reference RRBC, which reads rho/u from the macrovar output. */
void orc_rrbc(int kind, int lx, int ly, int lz, double* f, const double* fpost, const double* fx, const double* fy,
              const double* fz, const double* rho, const double* ux, const double* uy, const double* uz,
              int wall_x, int wall_y, int wall_z, int lid_face, const double* lid_u) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
    int walled[3] = {wall_x, wall_y, wall_z};
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                int nactive = 0;
                for (int face = 0; face < 6; ++face)
                    nactive += walled[face / 2] && on_face(face, ix, iy, iz, lx, ly, lz);
                if (nactive > 1) {
                    /* undo the bounce-back: periodic streaming from the upwind cell */
                    for (int ip = 0; ip < L.q; ++ip) {
                        int entering = 0;
                        for (int face = 0; face < 6; ++face)
                            entering |= walled[face / 2] && on_face(face, ix, iy, iz, lx, ly, lz) &&
                                        enters_through(&L, ip, face);
                        if (!entering) continue;
                        int sx = (ix - L.cx[ip] + lx) % lx, sy = (iy - L.cy[ip] + ly) % ly,
                            sz = (iz - L.cz[ip] + lz) % lz;
                        f[ip * N + idx] = fpost[ip * N + cidx(sx, sy, sz, lx, ly)];
                    }
                }
                for (int face = 0; face < 6; ++face) {
                    if (!walled[face / 2] || !on_face(face, ix, iy, iz, lx, ly, lz)) continue;
                    double fc[27], out[27];
                    if (nactive > 1) {
                        for (int ip = 0; ip < L.q; ++ip) {
                            if (!enters_through(&L, ip, face)) continue;
                            double v = fpost[L.opp[ip] * N + idx];
                            if (face == lid_face)
                                v += 6.0 * L.w[ip] * (L.cx[ip] * lid_u[0] + L.cy[ip] * lid_u[1] + L.cz[ip] * lid_u[2]);
                            f[ip * N + idx] = v;
                        }
                    }
                    for (int ip = 0; ip < L.q; ++ip) fc[ip] = f[ip * N + idx];
                    double r, u, v, w;
                    if (nactive == 1) {
                        r = rho[idx]; u = ux[idx]; v = uy[idx]; w = uz[idx];
                    } else {
                        double jx = 0, jy = 0, jz = 0;
                        r = 0;
                        for (int ip = 0; ip < L.q; ++ip) {
                            r += fc[ip]; jx += L.cx[ip] * fc[ip]; jy += L.cy[ip] * fc[ip]; jz += L.cz[ip] * fc[ip];
                        }
                        u = (jx + 0.5 * fx[idx]) / r; v = (jy + 0.5 * fy[idx]) / r; w = (jz + 0.5 * fz[idx]) / r;
                    }
                    rr_cell(&L, fc, r, u, v, w, fx[idx], fy[idx], fz[idx], 1.0, 1, out);
                    for (int ip = 0; ip < L.q; ++ip)
                        if (enters_through(&L, ip, face)) f[ip * N + idx] = out[ip];
                }
            }
}

/* FDBC, BC.cu:129-309: finite-difference wall reconstruction at the y walls
 * (the channel's alternative to RRBC, main.cu:76-77).  At each wall row
 * (iy = 0 and LY-1) the velocity is extrapolated from the two next rows to
 * the wall node (quadratic through u = 0 on the half-way wall,
 * BC.cu:149-161), the strain rate is taken by finite differences (one-sided
 * second order in y, central in x and z with periodic wrap, BC.cu:163-196),
 * the density from the ratio of the known populations' sum (c_y pointing into
 * the wall) to the same sum of the reconstruction at rho = 1
 * (BC.cu:203-265), and all populations are replaced by
 * feq + fneq(-rho tau_eff S) - Fbar feq at that density (BC.cu:268-308).
 * The reference kernel extrapolates and differentiates in one pass, so its x
 * and z differences read neighbours' wall velocities that other threads are
 * overwriting; here every wall velocity is extrapolated first (phase 1) and
 * the differences use the extrapolated values (phase 2).  Third order,
 * D3Q27 channel ordering; ly >= 4.  taut may be NULL. */
void orc_fdbc(int lx, int ly, int lz, double* f, const double* fx, const double* fy, const double* fz, double* ux,
              double* uy, double* uz, const double* taut, double tau) {
    orc_lat L;
    lat_build(ORC_D3Q27_CUDA, &L);
    const long N = (long)lx * ly * lz;
    /* phase 1: wall velocities (BC.cu:149-161) */
    for (int iz = 0; iz < lz; ++iz)
        for (int side = 0; side < 2; ++side)
            for (int ix = 0; ix < lx; ++ix) {
                int iy = side ? ly - 1 : 0, d = side ? -1 : 1;
                long idx = cidx(ix, iy, iz, lx, ly);
                long i1 = cidx(ix, iy + d, iz, lx, ly), i2 = cidx(ix, iy + 2 * d, iz, lx, ly);
                ux[idx] = 2.0 / 3.0 * ux[i1] - 1.0 / 5.0 * ux[i2];
                uy[idx] = 2.0 / 3.0 * uy[i1] - 1.0 / 5.0 * uy[i2];
                uz[idx] = 2.0 / 3.0 * uz[i1] - 1.0 / 5.0 * uz[i2];
            }
    /* phase 2: reconstruction */
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < lz; ++iz)
        for (int side = 0; side < 2; ++side)
            for (int ix = 0; ix < lx; ++ix) {
                int iy = side ? ly - 1 : 0, d = side ? -1 : 1;
                long idx = cidx(ix, iy, iz, lx, ly);
                long i1 = cidx(ix, iy + d, iz, lx, ly), i2 = cidx(ix, iy + 2 * d, iz, lx, ly);
                double tau_eff = tau + (taut ? taut[idx] : 0.0);
                /* BC.cu:167-180: (-3 u0 + 4 u1 - u2)/2 at the bottom, (3 u0 - 4 u1 + u2)/2 at the top */
                double sg = side ? -1.0 : 1.0;
                double dudy = sg * (-3.0 * ux[idx] + 4.0 * ux[i1] - ux[i2]) / 2.0;
                double dvdy = sg * (-3.0 * uy[idx] + 4.0 * uy[i1] - uy[i2]) / 2.0;
                double dwdy = sg * (-3.0 * uz[idx] + 4.0 * uz[i1] - uz[i2]) / 2.0;
                int ixp = (ix + 1) % lx, ixm = (ix - 1 + lx) % lx, izp = (iz + 1) % lz, izm = (iz - 1 + lz) % lz;
                long xp = cidx(ixp, iy, iz, lx, ly), xm = cidx(ixm, iy, iz, lx, ly);
                long zp = cidx(ix, iy, izp, lx, ly), zm = cidx(ix, iy, izm, lx, ly);
                double dudx = 0.5 * (ux[xp] - ux[xm]), dvdx = 0.5 * (uy[xp] - uy[xm]), dwdx = 0.5 * (uz[xp] - uz[xm]);
                double dudz = 0.5 * (ux[zp] - ux[zm]), dvdz = 0.5 * (uy[zp] - uy[zm]), dwdz = 0.5 * (uz[zp] - uz[zm]);
                double S110 = 0.5 * (dudy + dvdx), S101 = 0.5 * (dudz + dwdx), S011 = 0.5 * (dvdz + dwdy);
                double S200 = dudx, S020 = dvdy, S002 = dwdz;
                double u9 = ux[idx], v9 = uy[idx], w9 = uz[idx];
                double Fbar[27];
                fbar_cell(&L, u9, v9, w9, fx[idx], fy[idx], fz[idx], Fbar);
                double uu = u9 / ORC_CS, vv = v9 / ORC_CS, ww = w9 / ORC_CS;
                orc_a a;
                eq_coeffs(uu, vv, ww, &a);
                /* the reconstruction per unit density: feq(1) + fneq(-tau S) - Fbar feq(1), BC.cu:198-265 */
                orc_a n;
                n.a110 = -tau_eff * S110; n.a101 = -tau_eff * S101; n.a011 = -tau_eff * S011;
                n.a200 = -tau_eff * S200; n.a020 = -tau_eff * S020; n.a002 = -tau_eff * S002;
                n.a210 = n.a200 * vv + 2. * uu * n.a110;
                n.a201 = n.a200 * ww + 2. * uu * n.a101;
                n.a120 = n.a020 * uu + 2. * vv * n.a110;
                n.a021 = n.a020 * ww + 2. * vv * n.a011;
                n.a102 = n.a002 * uu + 2. * ww * n.a101;
                n.a012 = n.a002 * vv + 2. * ww * n.a011;
                n.a111 = n.a110 * ww + n.a101 * vv + n.a011 * uu;
                double g1[27], sum_ps = 0.0, sum_in = 0.0;
                for (int ip = 0; ip < 27; ++ip) {
                    double fir = L.H100[ip] * uu + L.H010[ip] * vv + L.H001[ip] * ww;
                    double feq1 = L.w[ip] * (1.0 + fir + 0.5 * (second_sum(&L, ip, &a) + third_sum(&L, ip, &a)));
                    double fneq1 = L.w[ip] * (second_sum(&L, ip, &n) + third_sum(&L, ip, &n));
                    g1[ip] = feq1 + fneq1 - Fbar[ip] * feq1;
                    int known = side ? L.cy[ip] > 0 : L.cy[ip] < 0;
                    if (known) { sum_ps += f[ip * N + idx]; sum_in += g1[ip]; }
                }
                /* every term of the reconstruction is proportional to rho (BC.cu:268-308) */
                double rho_l = sum_ps / sum_in;
                for (int ip = 0; ip < 27; ++ip) f[ip * N + idx] = rho_l * g1[ip];
            }
}

/* ------------------------------------------------------------------ */
/* WALE, forcing, initialisation                                       */
/* ------------------------------------------------------------------ */
/* LES_WALE, RR27V33_code_DNS_LES/LES.cu:4-98.  Central differences, periodic in
 * x and z; one-sided at the y walls (LES.cu:19-33).  guard != 0 returns 0
 * instead of 0/0 when every gradient vanishes. */
void orc_les_wale(int lx, int ly, int lz, const double* ux, const double* uy, const double* uz, double* taut,
                  double cw, int guard) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                double dudy, dvdy, dwdy, dudx, dvdx, dwdx, dudz, dvdz, dwdz;
                if (iy == 0) {
                    long p = cidx(ix, iy + 1, iz, lx, ly);
                    dudy = (3.0 * ux[idx] + ux[p]) / 3.0;
                    dvdy = (3.0 * uy[idx] + uy[p]) / 3.0;
                    dwdy = (3.0 * uz[idx] + uz[p]) / 3.0;
                } else if (iy == ly - 1) {
                    long m = cidx(ix, iy - 1, iz, lx, ly);
                    dudy = (-3.0 * ux[idx] - ux[m]) / 3.0;
                    dvdy = (-3.0 * uy[idx] - uy[m]) / 3.0;
                    dwdy = (-3.0 * uz[idx] - uz[m]) / 3.0;
                } else {
                    long p = cidx(ix, iy + 1, iz, lx, ly), m = cidx(ix, iy - 1, iz, lx, ly);
                    dudy = 0.5 * (ux[p] - ux[m]); dvdy = 0.5 * (uy[p] - uy[m]); dwdy = 0.5 * (uz[p] - uz[m]);
                }
                {
                    long p = cidx((ix + 1) % lx, iy, iz, lx, ly), m = cidx((ix - 1 + lx) % lx, iy, iz, lx, ly);
                    dudx = 0.5 * (ux[p] - ux[m]); dvdx = 0.5 * (uy[p] - uy[m]); dwdx = 0.5 * (uz[p] - uz[m]);
                }
                {
                    long p = cidx(ix, iy, (iz + 1) % lz, lx, ly), m = cidx(ix, iy, (iz - 1 + lz) % lz, lx, ly);
                    dudz = 0.5 * (ux[p] - ux[m]); dvdz = 0.5 * (uy[p] - uy[m]); dwdz = 0.5 * (uz[p] - uz[m]);
                }
                /* LES.cu:79-97 */
                double sijsq = (dudy + dvdx) * (dudy + dvdx) + (dudz + dwdx) * (dudz + dwdx) + (dvdz + dwdy) * (dvdz + dwdy);
                sijsq = 0.5 * sijsq + dudx * dudx + dvdy * dvdy + dwdz * dwdz;
                double s11d = (2. * dudx * dudx - dvdy * dvdy - dwdz * dwdz + dudy * dvdx + dudz * dwdx - 2. * dvdz * dwdy) / 3.;
                double s12d = 0.5 * (dudx * (dudy + dvdx) + dudy * dvdy + dvdx * dvdy + dvdz * dwdx + dudz * dwdy);
                double s13d = 0.5 * (dudx * (dudz + dwdx) + dudy * dvdz + dvdx * dwdy + dudz * dwdz + dwdx * dwdz);
                double s22d = (2. * dvdy * dvdy - dudx * dudx - dwdz * dwdz + dudy * dvdx + dvdz * dwdy - 2. * dudz * dwdx) / 3.;
                double s23d = 0.5 * (dvdy * (dvdz + dwdy) + dudz * dvdx + dudy * dwdx + dvdz * dwdz + dwdy * dwdz);
                double s33d = (2. * dwdz * dwdz - dudx * dudx - dvdy * dvdy + dudz * dwdx + dvdz * dwdy - 2. * dudy * dvdx) / 3.;
                double sd = s11d * s11d + s22d * s22d + s33d * s33d + 2.0 * (s12d * s12d + s13d * s13d + s23d * s23d);
                double den = pow(sijsq, 2.5) + pow(sd, 1.25);
                if (guard && den == 0.0)
                    taut[idx] = 0.0;
                else
                    taut[idx] = 3.0 * cw * cw * pow(sd, 1.5) / den;
            }
}

/* forcing, macrovar.cu:63-74 */
void orc_forcing(int lx, int ly, int lz, double* fx, double* fy, double* fz, double force_in_x) {
    long N = (long)lx * ly * lz;
    for (long i = 0; i < N; ++i) { fx[i] = force_in_x; fy[i] = 0.0; fz[i] = 0.0; }
}

/* forcing_GP, macrovar.cu:76-114.  The upper band uses integer division ix/LX
 * (always 0), as the reference does (macrovar.cu:106,109,112). */
void orc_forcing_gp(int lx, int ly, int lz, double* fx, double* fy, double* fz, double force_in_x, int istep,
                    double rand_phase) {
    const double pi2 = 2.0 * acos(-1.0);
    const double Tpd = 2000.0;
    const double LYq = ly / 4; /* integer division, macrovar.cu:85 */
    const int h0 = 2;
    const double beta9 = 3.0, gamma9 = 2.0;
    const double Amp0 = (100.0 * beta9 / ly) * sin(pi2 * istep / Tpd);
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                fx[idx] = force_in_x; fy[idx] = 0.0; fz[idx] = 0.0;
                if (iy > h0 && iy < LYq + h0) {
                    fx[idx] = force_in_x * (1.0 - Amp0 * lx / beta9 * sin(pi2 * (iy - h0) / LYq) *
                                                      sin(beta9 * pi2 * ix / lx) * cos(gamma9 * pi2 * iz / lz));
                    fy[idx] = force_in_x * Amp0 * 0.5 * LYq * (1 - cos(pi2 * (iy - h0) / LYq)) *
                              cos(beta9 * pi2 * ix / lx) * cos(gamma9 * pi2 * iz / lz);
                    fz[idx] = force_in_x * Amp0 * 0.5 * lz / gamma9 * sin(pi2 * (iy - h0) / LYq) *
                              cos(beta9 * pi2 * ix / lx) * sin(gamma9 * pi2 * iz / lz);
                } else if (iy > 0.75 * ly - h0 && iy < ly - h0) {
                    double xph = (double)(ix / lx) + rand_phase;
                    fx[idx] = force_in_x * (1.0 + Amp0 * lx / beta9 * sin(pi2 * (-ly + iy + h0 + LYq) / LYq) *
                                                      sin(beta9 * pi2 * xph) * cos(gamma9 * pi2 * iz / lz));
                    fy[idx] = -force_in_x * Amp0 * 0.5 * LYq * (1 - cos(pi2 * (-ly + iy + h0 + LYq) / LYq)) *
                              cos(beta9 * pi2 * xph) * cos(gamma9 * pi2 * iz / lz);
                    fz[idx] = -force_in_x * Amp0 * 0.5 * lz / gamma9 * sin(pi2 * (-ly + iy + h0 + LYq) / LYq) *
                              cos(beta9 * pi2 * xph) * sin(gamma9 * pi2 * iz / lz);
                }
            }
}

/* init_arrays, RR27V33_code_DNS_LES/initial.cu:268-325: rho = 1; u_x = viscous /
 * log10-law profile plus a sinusoidal perturbation; populations = rest
 * equilibrium shifted by half the force. */
void orc_init_arrays(int kind, int lx, int ly, int lz, double* f, double* rho, double* ux, double* uy, double* uz,
                     double* fx, double* fy, double* fz, double ustar, double ystar, double force_in_x) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
    const double pi2 = 2.0 * acos(-1.0);
    const double A9 = 0.1 * sqrt(2.0), alpha = 1.0, beta9 = 1.0, cc = 60.0;
    const double ccc1 = -lx / pi2 / alpha / ystar * A9 * ustar / cc / cc;
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                rho[idx] = 1.0;
                uy[idx] = 0.0; uz[idx] = 0.0;
                double z9 = pi2 * (iz + 0.5) / lz, x9 = pi2 * (ix + 0.5) / lx;
                double yplus = iy <= ly / 2 ? (iy + 0.5) / ystar : (ly - iy - 0.5) / ystar;
                double u = yplus <= 10.8 ? ustar * yplus : (log10(yplus) / 0.41 + 5.0) * ustar;
                u += ccc1 * yplus * sin(alpha * x9 + beta9 * z9);
                ux[idx] = u;
                fx[idx] = force_in_x; fy[idx] = 0.0; fz[idx] = 0.0;
                for (int ip = 0; ip < L.q; ++ip) {
                    double feq = L.w[ip] * rho[idx] * (1.0 - 1.5 * 0.0);
                    f[ip * N + idx] = feq - 0.5 * fx[idx] * L.cx[ip] / (1.0 / 3.0) * feq;
                }
            }
}

/* One time step in the channel driver's order, main.cu:50-84 (forcing arrays
 * are left as they are): [LES_WALE] -> collision_RR -> streaming -> BB ->
 * copy -> macrovar -> RRBC -> macrovar_BN. */
void orc_step_channel(int kind, int lx, int ly, int lz, int nsteps, double* f, double* ftmp, const double* fx,
                      const double* fy, const double* fz, double* rho, double* ux, double* uy, double* uz,
                      double* taut, double tau, int use_les, double cw, int les_guard, int wall_x, int wall_y,
                      int wall_z, int wall_rr, int lid_face, const double* lid_u) {
    const long N = (long)lx * ly * lz;
    const int Q = kind == ORC_D3Q19 ? 19 : 27;
    for (int s = 0; s < nsteps; ++s) {
        if (use_les) orc_les_wale(lx, ly, lz, ux, uy, uz, taut, cw, les_guard);
        orc_collision_rr(kind, lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, use_les ? taut : NULL, tau);
        orc_streaming(kind, lx, ly, lz, f, ftmp);
        orc_bb(kind, lx, ly, lz, f, ftmp, wall_x, wall_y, wall_z, lid_face, lid_u);
        /* after the copy f <- ftmp, ftmp keeps the post-collision populations */
        for (long i = 0; i < (long)Q * N; ++i) { double t = f[i]; f[i] = ftmp[i]; ftmp[i] = t; }
        orc_macrovar(kind, lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, 0, 0, 0, 0);
        if (wall_rr == 2 && wall_y) {
            /* FDBC in place of RRBC (main.cu:76-77), then macrovar_BN */
            orc_fdbc(lx, ly, lz, f, fx, fy, fz, ux, uy, uz, use_les ? taut : NULL, tau);
            orc_macrovar(kind, lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, 1, wall_x, wall_y, wall_z);
        } else if (wall_rr && (wall_x || wall_y || wall_z)) {
            orc_rrbc(kind, lx, ly, lz, f, ftmp, fx, fy, fz, rho, ux, uy, uz, wall_x, wall_y, wall_z, lid_face, lid_u);
            orc_macrovar(kind, lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, 1, wall_x, wall_y, wall_z);
        }
    }
}

/* The Laminar forced-channel solver, Laminar/main.f90:78-92 with
 * Laminar/collision.f90:8-175 (D3Q27, walls on the two x faces, periodic y
 * and z, uniform body force).  One step of its collision_RR: every node
 * collides with the macrovar state (collision.f90:30-61) and streams by swaps
 * with half-way bounce-back at the x walls (collision.f90:63-103) -- restated
 * here as push streaming + bounce-back, the same permutation -- then the
 * wall-incoming populations of the x = 1 and x = LX nodes (c_x > 0 and c_x < 0)
 * are replaced by feq + fneq - Fbar feq from the streamed populations, with
 * u = (j + F/2)/rho (collision.f90:107-172); then macrovar (main.f90:86).
 * stale_fbar != 0 keeps the reference's Fbar in the wall loops: the vector
 * left over from the last node of the collision sweep (x = LX, y = LY, z = LZ,
 * built from that node's pre-collision velocity, collision.f90:48-55); 0 uses
 * each wall node's own u (as the channel's RRBC and the GPU library do). */
void orc_step_laminar(int lx, int ly, int lz, int nsteps, double* f, double* ftmp, const double* fx,
                      const double* fy, const double* fz, double* rho, double* ux, double* uy, double* uz,
                      double tau, int stale_fbar) {
    orc_lat L;
    lat_build(ORC_D3Q27_CUDA, &L);
    const long N = (long)lx * ly * lz;
    for (int s = 0; s < nsteps; ++s) {
        double Fst[27];
        const long last = N - 1;
        fbar_cell(&L, ux[last], uy[last], uz[last], fx[last], fy[last], fz[last], Fst);
        orc_collision_rr(ORC_D3Q27_CUDA, lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, NULL, tau);
        orc_streaming(ORC_D3Q27_CUDA, lx, ly, lz, f, ftmp);
        orc_bb(ORC_D3Q27_CUDA, lx, ly, lz, f, ftmp, 1, 0, 0, -1, NULL);
        for (long i = 0; i < 27 * N; ++i) { double t = f[i]; f[i] = ftmp[i]; ftmp[i] = t; }
        for (int side = 0; side < 2; ++side) {
            const int ix = side == 0 ? 0 : lx - 1;
            for (int iz = 0; iz < lz; ++iz)
                for (int iy = 0; iy < ly; ++iy) {
                    const long idx = cidx(ix, iy, iz, lx, ly);
                    double fc[27], out[27], Fb[27];
                    double r = 0, jx = 0, jy = 0, jz = 0;
                    for (int ip = 0; ip < 27; ++ip) {
                        fc[ip] = f[ip * N + idx];
                        r += fc[ip]; jx += L.cx[ip] * fc[ip]; jy += L.cy[ip] * fc[ip]; jz += L.cz[ip] * fc[ip];
                    }
                    const double u = (jx + fx[idx] / 2.) / r, v = (jy + fy[idx] / 2.) / r, w = (jz + fz[idx] / 2.) / r;
                    if (stale_fbar)
                        for (int ip = 0; ip < 27; ++ip) Fb[ip] = Fst[ip];
                    else
                        fbar_cell(&L, u, v, w, fx[idx], fy[idx], fz[idx], Fb);
                    rr_cell_fbar(&L, fc, r, u, v, w, Fb, 1.0, 1, out);
                    for (int ip = 0; ip < 27; ++ip)
                        if ((side == 0 && L.cx[ip] > 0) || (side == 1 && L.cx[ip] < 0)) f[ip * N + idx] = out[ip];
                }
        }
        orc_macrovar(ORC_D3Q27_CUDA, lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, 0, 0, 0, 0);
    }
}

/* ------------------------------------------------------------------ */
/* The BGK baseline solver (BGK/): incompressible D3Q27 BGK on shifted  */
/* populations, density deviation rho and u = j + F/2 (no division)     */
/* ------------------------------------------------------------------ */
/* collision_BGK, BGK/collision.cu:3-33 (in place): with eu = c.u/cs^2 and
 * uv = |u|^2/cs^2, feq = w (rho + eu + (eu^2 - uv)/2), feq0 the same with
 * rho = 1, f += -(f - feq)/tau + (1 - 1/(2 tau)) (c - u).F / cs^2 feq0. */
void orc_collision_bgk(int lx, int ly, int lz, double* f, const double* fx, const double* fy, const double* fz,
                       const double* rho, const double* ux, const double* uy, const double* uz, double tau) {
    orc_lat L;
    lat_build(ORC_D3Q27_CUDA, &L);
    const long N = (long)lx * ly * lz;
    const double RT = 1.0 / 3.0;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; ++idx) {
        const double u = ux[idx], v = uy[idx], w = uz[idx];
        const double G3 = u * fx[idx] + v * fy[idx] + w * fz[idx];
        for (int ip = 0; ip < 27; ++ip) {
            const double eu = (L.cx[ip] * u + L.cy[ip] * v + L.cz[ip] * w) / RT;
            const double uv = (u * u + v * v + w * w) / RT;
            const double feq = L.w[ip] * (rho[idx] + eu + 0.5 * (eu * eu - uv));
            const double G1 = L.cx[ip] * fx[idx] + L.cy[ip] * fy[idx] + L.cz[ip] * fz[idx];
            double fchange = -(f[ip * N + idx] - feq) / tau;
            const double feq0 = L.w[ip] * (1.0 + eu + 0.5 * (eu * eu - uv));
            fchange += (1.0 - 0.5 / tau) * (G1 - G3) / RT * feq0;
            f[ip * N + idx] += fchange;
        }
    }
}

/* macrovar of the BGK solver, BGK/macrovar.cu:4-31: rho = sum f (the density
 * deviation), u = sum c f + F/2. */
void orc_macrovar_bgk(int lx, int ly, int lz, const double* f, const double* fx, const double* fy, const double* fz,
                      double* rho, double* ux, double* uy, double* uz) {
    orc_lat L;
    lat_build(ORC_D3Q27_CUDA, &L);
    const long N = (long)lx * ly * lz;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; ++idx) {
        double r = 0, jx = 0, jy = 0, jz = 0;
        for (int ip = 0; ip < 27; ++ip) {
            const double v = f[ip * N + idx];
            r += v; jx += L.cx[ip] * v; jy += L.cy[ip] * v; jz += L.cz[ip] * v;
        }
        rho[idx] = r;
        ux[idx] = jx + 0.5 * fx[idx];
        uy[idx] = jy + 0.5 * fy[idx];
        uz[idx] = jz + 0.5 * fz[idx];
    }
}

/* One step of the BGK solver's loop, BGK/main.cu:49-70 (the force arrays as
 * they are): collision_BGK -> streaming (periodic push, BGK/streaming.cu:3-29)
 * -> streaming2 (half-way bounce-back at the y walls, :31-63) -> copy ->
 * macrovar.  walls: as orc_bb (the reference has y walls). */
void orc_step_bgk(int lx, int ly, int lz, int nsteps, double* f, double* ftmp, const double* fx, const double* fy,
                  const double* fz, double* rho, double* ux, double* uy, double* uz, double tau, int wall_x,
                  int wall_y, int wall_z) {
    const long N = (long)lx * ly * lz;
    for (int s = 0; s < nsteps; ++s) {
        orc_collision_bgk(lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz, tau);
        orc_streaming(ORC_D3Q27_CUDA, lx, ly, lz, f, ftmp);
        orc_bb(ORC_D3Q27_CUDA, lx, ly, lz, f, ftmp, wall_x, wall_y, wall_z, -1, NULL);
        for (long i = 0; i < 27 * N; ++i) { double t = f[i]; f[i] = ftmp[i]; ftmp[i] = t; }
        orc_macrovar_bgk(lx, ly, lz, f, fx, fy, fz, rho, ux, uy, uz);
    }
}

/* ------------------------------------------------------------------ */
/* Diagnostics                                                         */
/* ------------------------------------------------------------------ */
/* diag_flow sums, output.cu:24-35: sum(u), sum(v), sum(w), sum(rho) in
 * index order, plus mass-excess, momentum and kinetic energy sums. */
void orc_diag(long n, const double* rho, const double* ux, const double* uy, const double* uz, double* out /* 9 */) {
    double s[9] = {0};
    for (long i = 0; i < n; ++i) {
        s[0] += ux[i]; s[1] += uy[i]; s[2] += uz[i]; s[3] += rho[i];
        s[4] += rho[i] - 1.0;
        s[5] += rho[i] * ux[i]; s[6] += rho[i] * uy[i]; s[7] += rho[i] * uz[i];
        s[8] += 0.5 * rho[i] * (ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]);
    }
    memcpy(out, s, sizeof s);
}

/* output_profile per-plane statistics, output.cu:182-215: for each iy the
 * plane means of u, v, w, tau_t and the six covariances of the fluctuations.
 * out[iy*10 + k], k = <u>,<v>,<w>,<uu>,<vv>,<ww>,<uv>,<uw>,<vw>,<taut>. */
void orc_profiles(int lx, int ly, int lz, const double* ux, const double* uy, const double* uz, const double* taut,
                  double* out) {
    for (int iy = 0; iy < ly; ++iy) {
        double su = 0, sv = 0, sw = 0, st = 0;
        long count = 0;
        for (int ix = 0; ix < lx; ++ix)
            for (int iz = 0; iz < lz; ++iz) {
                long idx = cidx(ix, iy, iz, lx, ly);
                su += ux[idx]; sv += uy[idx]; sw += uz[idx]; st += taut ? taut[idx] : 0.0;
                ++count;
            }
        double au = su / count, av = sv / count, aw = sw / count, at = st / count;
        double s_uu = 0, s_vv = 0, s_ww = 0, s_uv = 0, s_uw = 0, s_vw = 0;
        for (int ix = 0; ix < lx; ++ix)
            for (int iz = 0; iz < lz; ++iz) {
                long idx = cidx(ix, iy, iz, lx, ly);
                double ur = ux[idx] - au, vr = uy[idx] - av, wr = uz[idx] - aw;
                s_uu += ur * ur; s_vv += vr * vr; s_ww += wr * wr;
                s_uv += ur * vr; s_uw += ur * wr; s_vw += vr * wr;
            }
        double* o = out + 10 * iy;
        o[0] = au; o[1] = av; o[2] = aw;
        o[3] = s_uu / count; o[4] = s_vv / count; o[5] = s_ww / count;
        o[6] = s_uv / count; o[7] = s_uw / count; o[8] = s_vw / count; o[9] = at;
    }
}

/* ------------------------------------------------------------------ */
/* Seeded configuration initialisers (restating the library's own       */
/* counter-based generator and profiles so the GPU init can be checked)  */
/* ------------------------------------------------------------------ */
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
/* uniform in [0,1) from (seed, stream, counter) */
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t counter) {
    uint64_t h = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ull + counter));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

/* Mean profile U+(y+): y+ in the viscous layer, ln(y+)/0.41 + 5 above 10.8. */
static double uplus(double yplus) { return yplus <= 10.8 ? yplus : log(yplus) / 0.41 + 5.0; }

/* Channel initial field (the library's rrlbm_init_channel): rho = 1,
 * u_x = u_tau U+(y+) with y+ from the nearer half-way wall, plus streaks
 * 0.1 U(y) sin(2 pi kz z/LZ) cos(2 pi kx x/LX + phi) and uniform noise
 * amp_noise*U_b in [-1,1) on all three components. */
void orc_init_channel_fields(int lx, int ly, int lz, double ustar, double visc, uint64_t seed, double amp_streak,
                             double amp_noise, double* rho, double* ux, double* uy, double* uz) {
    const double pi2 = 2.0 * acos(-1.0);
    int kx = 2 + (int)(splitmix64(seed + 1) % 3), kz = 2 + (int)(splitmix64(seed + 2) % 3);
    double phi = pi2 * orc_uniform(seed, 3, 0);
    double ystar = visc / ustar;
    double ub = 0;
    for (int iy = 0; iy < ly; ++iy) {
        double d = iy < ly / 2 ? iy + 0.5 : ly - iy - 0.5;
        ub += ustar * uplus(d / ystar);
    }
    ub /= ly;
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                double d = iy < ly / 2 ? iy + 0.5 : ly - iy - 0.5;
                double U = ustar * uplus(d / ystar);
                double s = amp_streak * U * sin(pi2 * kz * (iz + 0.5) / lz) * cos(pi2 * kx * (ix + 0.5) / lx + phi);
                uint64_t c = (uint64_t)idx;
                rho[idx] = 1.0;
                ux[idx] = U + s + amp_noise * ub * (2.0 * orc_uniform(seed, 10, c) - 1.0);
                uy[idx] = amp_noise * ub * (2.0 * orc_uniform(seed, 11, c) - 1.0);
                uz[idx] = amp_noise * ub * (2.0 * orc_uniform(seed, 12, c) - 1.0);
            }
}

/* Populations = RR equilibrium of (rho, u) per cell. */
void orc_init_equilibrium(int kind, int lx, int ly, int lz, const double* rho, const double* ux, const double* uy,
                          const double* uz, double* f) {
    orc_lat L;
    lat_build(kind, &L);
    const long N = (long)lx * ly * lz;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; ++idx) {
        double feq[27];
        feq_cell(&L, rho[idx], ux[idx], uy[idx], uz[idx], feq);
        for (int ip = 0; ip < L.q; ++ip) f[ip * N + idx] = feq[ip];
    }
}

/* ------------------------------------------------------------------ */
/* Fortran Taylor-Green solver, Taylor_Green/D3Q27RR.f90 (and D3Q19RR.f90) */
/* ------------------------------------------------------------------ */
/* The Fortran declares u~, the a/ane coefficients, sec/thir/eu and the Hermite
 * tables as default REAL (D3Q27RR.f90:45-61) while f, rho, u are REAL(8).
 * real4 != 0 reproduces that mixed precision (float where the Fortran has
 * REAL, double where it has REAL(8), with Fortran's promotion rules);
 * real4 == 0 runs everything in double. */
typedef struct {
    int q, L, LZ;
    int ix[27], iy[27], iz[27];
    double wwp[27];
    double H[16][27]; /* as double; rounded to float when used in real4 mode */
} tg_ctx;

#define TGIDX(ip, i, j, k, q, L) ((ip) + (long)(q) * ((i) + (long)(L) * ((j) + (long)(L) * (k))))
#define FIDX(i, j, k, L) ((i) + (long)(L) * ((j) + (long)(L) * (k)))

static void tg_setup(int kind, int L, int LZ, tg_ctx* t) {
    orc_lat lat;
    lat_build(kind == ORC_D3Q19 ? ORC_D3Q19 : ORC_D3Q27_FORTRAN, &lat);
    t->q = lat.q; t->L = L; t->LZ = LZ;
    for (int i = 0; i < lat.q; ++i) {
        t->ix[i] = lat.cx[i]; t->iy[i] = lat.cy[i]; t->iz[i] = lat.cz[i]; t->wwp[i] = lat.w[i];
        double* src[16] = {lat.H100, lat.H010, lat.H001, lat.H110, lat.H101, lat.H011, lat.H200, lat.H020,
                           lat.H002, lat.H210, lat.H201, lat.H021, lat.H120, lat.H102, lat.H012, lat.H111};
        for (int k = 0; k < 16; ++k) t->H[k][i] = src[k][i];
    }
}
enum { T100, T010, T001, T110, T101, T011, T200, T020, T002, T210, T201, T021, T120, T102, T012, T111 };

/* feq (first) and fneq (second) sums for one direction, in float
 * (REAL) arithmetic; a[] and n[] in the order 110 101 011 200 020 002 210 201
 * 120 021 102 012 111 (D3Q27RR.f90:216-222,253-259; D3Q19RR.f90:230-235,274-279). */
static float tg_sec_f(const tg_ctx* t, int ip, const float* a) {
    const float H110 = (float)t->H[T110][ip], H101 = (float)t->H[T101][ip], H011 = (float)t->H[T011][ip];
    const float H200 = (float)t->H[T200][ip], H020 = (float)t->H[T020][ip], H002 = (float)t->H[T002][ip];
    return 2.f * (H110 * a[0] + H101 * a[1] + H011 * a[2]) + H200 * a[3] + H020 * a[4] + H002 * a[5];
}
static float tg_thir_f(const tg_ctx* t, int ip, const float* a) {
    if (t->q == 27) {
        const float H210 = (float)t->H[T210][ip], H201 = (float)t->H[T201][ip], H021 = (float)t->H[T021][ip];
        const float H120 = (float)t->H[T120][ip], H102 = (float)t->H[T102][ip], H012 = (float)t->H[T012][ip];
        const float H111 = (float)t->H[T111][ip];
        /* order of D3Q27RR.f90:219-221: 210 201 120 021 102 012 111 */
        return H210 * a[6] + H201 * a[7] + H120 * a[8] + H021 * a[9] + H102 * a[10] + H012 * a[11] + 2.f * H111 * a[12];
    }
    /* D3Q19RR.f90:131-136 builds these as (real8 expr) +/- (real8 expr) -> REAL */
    const float Hp1 = (float)(t->H[T210][ip] + t->H[T012][ip]), Hm1 = (float)(t->H[T210][ip] - t->H[T012][ip]);
    const float Hp2 = (float)(t->H[T102][ip] + t->H[T120][ip]), Hm2 = (float)(t->H[T102][ip] - t->H[T120][ip]);
    const float Hp3 = (float)(t->H[T021][ip] + t->H[T201][ip]), Hm3 = (float)(t->H[T021][ip] - t->H[T201][ip]);
    const float ap1 = a[6] + a[11], am1 = a[6] - a[11];
    const float ap2 = a[10] + a[8], am2 = a[10] - a[8];
    const float ap3 = a[9] + a[7], am3 = a[9] - a[7];
    /* left to right in REAL until the first /3.d0 promotes the sum to double,
     * D3Q19RR.f90:230-235 */
    const float plus = (Hp1 * ap1 + Hp2 * ap2) + Hp3 * ap3;
    double s = (double)plus + (double)(Hm1 * am1) / 3.0;
    s = s + (double)(Hm2 * am2) / 3.0;
    s = s + (double)(Hm3 * am3) / 3.0;
    return (float)s;
}
static double tg_sec_d(const tg_ctx* t, int ip, const double* a) {
    return 2.0 * (t->H[T110][ip] * a[0] + t->H[T101][ip] * a[1] + t->H[T011][ip] * a[2]) + t->H[T200][ip] * a[3] +
           t->H[T020][ip] * a[4] + t->H[T002][ip] * a[5];
}
static double tg_thir_d(const tg_ctx* t, int ip, const double* a) {
    if (t->q == 27)
        return t->H[T210][ip] * a[6] + t->H[T201][ip] * a[7] + t->H[T120][ip] * a[8] + t->H[T021][ip] * a[9] +
               t->H[T102][ip] * a[10] + t->H[T012][ip] * a[11] + 2.0 * t->H[T111][ip] * a[12];
    double Hp1 = t->H[T210][ip] + t->H[T012][ip], Hm1 = t->H[T210][ip] - t->H[T012][ip];
    double Hp2 = t->H[T102][ip] + t->H[T120][ip], Hm2 = t->H[T102][ip] - t->H[T120][ip];
    double Hp3 = t->H[T021][ip] + t->H[T201][ip], Hm3 = t->H[T021][ip] - t->H[T201][ip];
    return Hp1 * (a[6] + a[11]) + Hp2 * (a[10] + a[8]) + Hp3 * (a[9] + a[7]) + Hm1 * (a[6] - a[11]) / 3.0 +
           Hm2 * (a[10] - a[8]) / 3.0 + Hm3 * (a[9] - a[7]) / 3.0;
}

/* feq of one cell, D3Q27RR.f90:204-223 / :327-350 */
static void tg_feq(const tg_ctx* t, int real4, double u9, double v9, double w9, double rho9, double* feq9) {
    const double cs = 1.0 / sqrt(3.0);
    if (real4) {
        float uu = (float)(u9 / cs), vv = (float)(v9 / cs), ww = (float)(w9 / cs);
        float a[13];
        a[0] = uu * vv; a[1] = uu * ww; a[2] = vv * ww; a[3] = uu * uu; a[4] = vv * vv; a[5] = ww * ww;
        a[6] = a[3] * vv; a[7] = a[3] * ww; a[8] = a[4] * uu; a[9] = a[4] * ww; a[10] = a[5] * uu; a[11] = a[5] * vv;
        a[12] = a[0] * ww;
        for (int ip = 0; ip < t->q; ++ip) {
            float eu = (float)t->H[T100][ip] * uu + (float)t->H[T010][ip] * vv + (float)t->H[T001][ip] * ww;
            float sec = tg_sec_f(t, ip, a), thir = tg_thir_f(t, ip, a);
            float fe = 1.f + eu + 0.5f * (sec + thir);
            feq9[ip] = t->wwp[ip] * rho9 * (double)fe;
        }
    } else {
        double uu = u9 / cs, vv = v9 / cs, ww = w9 / cs;
        double a[13];
        a[0] = uu * vv; a[1] = uu * ww; a[2] = vv * ww; a[3] = uu * uu; a[4] = vv * vv; a[5] = ww * ww;
        a[6] = a[3] * vv; a[7] = a[3] * ww; a[8] = a[4] * uu; a[9] = a[4] * ww; a[10] = a[5] * uu; a[11] = a[5] * vv;
        a[12] = a[0] * ww;
        for (int ip = 0; ip < t->q; ++ip) {
            double eu = t->H[T100][ip] * uu + t->H[T010][ip] * vv + t->H[T001][ip] * ww;
            feq9[ip] = t->wwp[ip] * rho9 * (1.0 + eu + 0.5 * (tg_sec_d(t, ip, a) + tg_thir_d(t, ip, a)));
        }
    }
}

/* fneq of one cell from the second-order non-equilibrium moments n2[6]
 * (110 101 011 200 020 002), via the recursion D3Q27RR.f90:375-390. */
static void tg_fneq(const tg_ctx* t, int real4, double u9, double v9, double w9, const double* n2in,
                    double* fneq9) {
    const double cs = 1.0 / sqrt(3.0);
    if (real4) {
        float uu = (float)(u9 / cs), vv = (float)(v9 / cs), ww = (float)(w9 / cs);
        float n[13];
        for (int k = 0; k < 6; ++k) n[k] = (float)n2in[k];
        n[6] = n[3] * vv + 2.f * uu * n[0];  /* ane210 */
        n[7] = n[3] * ww + 2.f * uu * n[1];  /* ane201 */
        n[8] = n[4] * uu + 2.f * vv * n[0];  /* ane120 */
        n[9] = n[4] * ww + 2.f * vv * n[2];  /* ane021 */
        n[10] = n[5] * uu + 2.f * ww * n[1]; /* ane102 */
        n[11] = n[5] * vv + 2.f * ww * n[2]; /* ane012 */
        n[12] = n[0] * ww + n[1] * vv + n[2] * uu;
        for (int ip = 0; ip < t->q; ++ip) {
            float sec = tg_sec_f(t, ip, n), thir = tg_thir_f(t, ip, n);
            fneq9[ip] = t->wwp[ip] * 0.5 * (double)(sec + thir);
        }
    } else {
        double uu = u9 / cs, vv = v9 / cs, ww = w9 / cs;
        double n[13];
        for (int k = 0; k < 6; ++k) n[k] = n2in[k];
        n[6] = n[3] * vv + 2. * uu * n[0]; n[7] = n[3] * ww + 2. * uu * n[1];
        n[8] = n[4] * uu + 2. * vv * n[0]; n[9] = n[4] * ww + 2. * vv * n[2];
        n[10] = n[5] * uu + 2. * ww * n[1]; n[11] = n[5] * vv + 2. * ww * n[2];
        n[12] = n[0] * ww + n[1] * vv + n[2] * uu;
        for (int ip = 0; ip < t->q; ++ip)
            fneq9[ip] = t->wwp[ip] * 0.5 * (tg_sec_d(t, ip, n) + tg_thir_d(t, ip, n));
    }
}

/* Initial state, D3Q27RR.f90:134-264: analytic TG fields and
 * f = feq + fneq(4th-order central-difference gradients). */
void orc_tg_ref_init(int kind, int L, int LZ, int real4, double* f, double* rho, double* u, double* v, double* w) {
    tg_ctx t;
    tg_setup(kind, L, LZ, &t);
    const double pi = 3.14159265358979323846, U0 = 1.0 * (1.0 / sqrt(3.0)) / 100.0, kv = 2.0;
    const double cs = 1.0 / sqrt(3.0), cs2 = 1.0 / 3.0;
    const double Re0 = 200.0 / cs, visc = U0 * L / Re0, tau = visc / cs2 + 0.5;
    for (int k = 0; k < LZ; ++k)
        for (int j = 0; j < L; ++j)
            for (int i = 0; i < L; ++i) {
                double x = (i + 1) - 0.5, y = (j + 1) - 0.5;
                long c = FIDX(i, j, k, L);
                u[c] = -U0 * cos(kv * pi * x / L) * sin(kv * pi * y / L);
                v[c] = U0 * cos(kv * pi * y / L) * sin(kv * pi * x / L);
                w[c] = 0.0;
                rho[c] = 1.0 - U0 * U0 / 2 * cos(kv * pi / L * (x - y)) * cos(kv * pi / L * (x + y)) / cs2;
            }
    for (int k = 0; k < LZ; ++k)
        for (int j = 0; j < L; ++j)
            for (int i = 0; i < L; ++i) {
                long c = FIDX(i, j, k, L);
                int ip1 = (i + 1) % L, ip2 = (i + 2) % L, im1 = (i - 1 + L) % L, im2 = (i - 2 + L) % L;
                int jp1 = (j + 1) % L, jp2 = (j + 2) % L, jm1 = (j - 1 + L) % L, jm2 = (j - 2 + L) % L;
                double dudx = -(u[FIDX(ip2, j, k, L)] - 8.0 * u[FIDX(ip1, j, k, L)] + 8.0 * u[FIDX(im1, j, k, L)] - u[FIDX(im2, j, k, L)]) / 12.0;
                double dudy = -(u[FIDX(i, jp2, k, L)] - 8.0 * u[FIDX(i, jp1, k, L)] + 8.0 * u[FIDX(i, jm1, k, L)] - u[FIDX(i, jm2, k, L)]) / 12.0;
                double dvdx = -(v[FIDX(ip2, j, k, L)] - 8.0 * v[FIDX(ip1, j, k, L)] + 8.0 * v[FIDX(im1, j, k, L)] - v[FIDX(im2, j, k, L)]) / 12.0;
                double dvdy = -(v[FIDX(i, jp2, k, L)] - 8.0 * v[FIDX(i, jp1, k, L)] + 8.0 * v[FIDX(i, jm1, k, L)] - v[FIDX(i, jm2, k, L)]) / 12.0;
                double feq9[27], fneq9[27], n2[6];
                tg_feq(&t, real4, u[c], v[c], w[c], rho[c], feq9);
                /* D3Q27RR.f90:237-242 */
                n2[0] = -tau * rho[c] * (dudy + dvdx);
                n2[1] = 0.0; n2[2] = 0.0;
                n2[3] = -tau * rho[c] * 2.0 * dudx;
                n2[4] = -tau * rho[c] * 2.0 * dvdy;
                n2[5] = 0.0;
                tg_fneq(&t, real4, u[c], v[c], w[c], n2, fneq9);
                for (int ip = 0; ip < t.q; ++ip) f[TGIDX(ip, i, j, k, t.q, L)] = feq9[ip] + fneq9[ip];
            }
}

/* nsteps iterations of D3Q27RR.f90:319-442: RR collision without force
 * (using rho,u of the previous macrovar), pull streaming, macrovar. */
void orc_tg_ref_run(int kind, int L, int LZ, int real4, int nsteps, double tau, double* f, double* rho, double* u,
                    double* v, double* w) {
    tg_ctx t;
    tg_setup(kind, L, LZ, &t);
    const long NC = (long)L * L * LZ;
    double* fpost = (double*)malloc(sizeof(double) * t.q * NC);
    const double cs = 1.0 / sqrt(3.0);
    for (int it = 0; it < nsteps; ++it) {
#pragma omp parallel for schedule(static)
        for (long c = 0; c < NC; ++c) {
            double u9 = u[c], v9 = v[c], w9 = w[c], rho9 = rho[c];
            double feq9[27], fneq9[27], n2[6];
            tg_feq(&t, real4, u9, v9, w9, rho9, feq9);
            const double* fc = f + (long)t.q * c;
            if (real4) {
                float uu = (float)(u9 / cs), vv = (float)(v9 / cs), ww = (float)(w9 / cs);
                float a110 = (float)(rho9 * uu * vv), a101 = (float)(rho9 * uu * ww), a011 = (float)(rho9 * vv * ww);
                float a200 = (float)(rho9 * uu * uu), a020 = (float)(rho9 * vv * vv), a002 = (float)(rho9 * ww * ww);
                float e110 = 0.f, e101 = 0.f, e011 = 0.f, e200 = 0.f, e020 = 0.f, e002 = 0.f;
                for (int ip = 0; ip < t.q; ++ip) {
                    double f9 = fc[ip];
                    e110 = (float)(e110 + (float)t.H[T110][ip] * f9);
                    e101 = (float)(e101 + (float)t.H[T101][ip] * f9);
                    e011 = (float)(e011 + (float)t.H[T011][ip] * f9);
                    e200 = (float)(e200 + (float)t.H[T200][ip] * f9);
                    e020 = (float)(e020 + (float)t.H[T020][ip] * f9);
                    e002 = (float)(e002 + (float)t.H[T002][ip] * f9);
                }
                n2[0] = (float)(e110 - a110); n2[1] = (float)(e101 - a101); n2[2] = (float)(e011 - a011);
                n2[3] = (float)(e200 - a200); n2[4] = (float)(e020 - a020); n2[5] = (float)(e002 - a002);
            } else {
                double uu = u9 / cs, vv = v9 / cs, ww = w9 / cs;
                double e[6] = {0, 0, 0, 0, 0, 0};
                for (int ip = 0; ip < t.q; ++ip) {
                    double f9 = fc[ip];
                    e[0] += t.H[T110][ip] * f9; e[1] += t.H[T101][ip] * f9; e[2] += t.H[T011][ip] * f9;
                    e[3] += t.H[T200][ip] * f9; e[4] += t.H[T020][ip] * f9; e[5] += t.H[T002][ip] * f9;
                }
                n2[0] = e[0] - rho9 * uu * vv; n2[1] = e[1] - rho9 * uu * ww; n2[2] = e[2] - rho9 * vv * ww;
                n2[3] = e[3] - rho9 * uu * uu; n2[4] = e[4] - rho9 * vv * vv; n2[5] = e[5] - rho9 * ww * ww;
            }
            tg_fneq(&t, real4, u9, v9, w9, n2, fneq9);
            /* D3Q27RR.f90:394: (1. - 1./tau) with a REAL literal 1. */
            double om = real4 ? (double)1.f - 1.0 / tau : 1.0 - 1.0 / tau;
            for (int ip = 0; ip < t.q; ++ip) fpost[(long)t.q * c + ip] = feq9[ip] + om * fneq9[ip];
        }
        /* pull streaming, D3Q27RR.f90:400-419 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < LZ; ++k)
            for (int j = 0; j < L; ++j)
                for (int i = 0; i < L; ++i) {
                    for (int ip = 0; ip < t.q; ++ip) {
                        int ipre = (i - t.ix[ip] + L) % L, jpre = (j - t.iy[ip] + L) % L, kpre = (k - t.iz[ip] + LZ) % LZ;
                        f[TGIDX(ip, i, j, k, t.q, L)] = fpost[TGIDX(ip, ipre, jpre, kpre, t.q, L)];
                    }
                }
        /* macrovar, D3Q27RR.f90:421-442 */
#pragma omp parallel for schedule(static)
        for (long c = 0; c < NC; ++c) {
            double r = 0, a = 0, b = 0, d = 0;
            for (int ip = 0; ip < t.q; ++ip) {
                double fv = f[(long)t.q * c + ip];
                r += fv; a += fv * t.ix[ip]; b += fv * t.iy[ip]; d += fv * t.iz[ip];
            }
            rho[c] = r; u[c] = a / r; v[c] = b / r; w[c] = d / r;
        }
    }
    free(fpost);
}

/* Taylor-Green initial fields of the configurations (the library's
 * rrlbm_init_taylor_green_2d / _3d), cell centres x = ix + 1/2:
 *   2D: u = u0 sin(kx x) cos(ky y), v = -u0 cos(kx x) sin(ky y), w = 0,
 *       rho = 1 + u0^2/(4 cs^2) (cos 2kx x + cos 2ky y), k = 2 pi / L;
 *   3D: on [0, 2 pi)^3, u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z,
 *       w = 0, rho = 1 + (u0^2/16)(cos 2x + cos 2y)(cos 2z + 2)/cs^2. */
void orc_init_tg_fields(int three_d, int lx, int ly, int lz, double u0, double* rho, double* ux, double* uy,
                        double* uz) {
    const double pi2 = 2.0 * acos(-1.0);
    for (int iz = 0; iz < lz; ++iz)
        for (int iy = 0; iy < ly; ++iy)
            for (int ix = 0; ix < lx; ++ix) {
                long idx = cidx(ix, iy, iz, lx, ly);
                if (!three_d) {
                    double kx = pi2 / lx, ky = pi2 / ly, x = ix + 0.5, y = iy + 0.5;
                    ux[idx] = u0 * sin(kx * x) * cos(ky * y);
                    uy[idx] = -u0 * cos(kx * x) * sin(ky * y);
                    uz[idx] = 0.0;
                    rho[idx] = 1.0 + u0 * u0 / (4.0 / 3.0) * (cos(2.0 * kx * x) + cos(2.0 * ky * y));
                } else {
                    double x = pi2 * (ix + 0.5) / lx, y = pi2 * (iy + 0.5) / ly, z = pi2 * (iz + 0.5) / lz;
                    ux[idx] = u0 * sin(x) * cos(y) * cos(z);
                    uy[idx] = -u0 * cos(x) * sin(y) * cos(z);
                    uz[idx] = 0.0;
                    rho[idx] = 1.0 + 3.0 * (u0 * u0 / 16.0) * (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0);
                }
            }
}
