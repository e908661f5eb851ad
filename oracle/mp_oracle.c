/*
 * ORACLE -- test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do, to check the GPU results and to time a CPU baseline.
 *
 * Plain-C, FP64, OpenMP restatement of the reference's Multipole path:
 *   table build   gpu_potential.py:183-626   (_compute_invPhi0,
 *                 _compute_outer_extrap, _solve_quintic_d2,
 *                 _build_multipole_data)
 *   evaluation    _multipole_potential_kernel.cu:100-958
 *                 (quintic_eval, multipole_eval_device<DO_GRAD>,
 *                  multipole_hess_kernel, multipole_density_kernel)
 *
 * It keeps the reference's formulation on purpose -- un-normalised Legendre
 * recurrence with the running normalisation factor, atan2 + rolling
 * cos/sin(m phi), one quintic Horner evaluation per active term -- so that it
 * is an independent check of the product's restructured kernels.
 *
 * Extrapolation laws: the reference's Lambert-W branch choice
 * (gpu_potential.py:226-229, 315-316) is the trivial root s = v; this file
 * evaluates its formulas in exact arithmetic, i.e. the s == v "log" laws
 *   inner  Phi = W + U ln(r/r0)               (v = 0)
 *   outer  Phi = (W + U ln(r/rN)) (r/rN)^-1   (v = -1)
 * unless the inner s=2 override applies (gpu_potential.py:245-251).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_LMAX 32

typedef struct {
    int n_int, n_lm, lmax, log_scaling;
    int inner_log, outer_log;
    double logr_min, dlogr, inv_dlogr, invPhi0;
    double inner_s, inner_U, inner_W;
    double outer_s, outer_U, outer_W;
    int *lm_l, *lm_m;          /* n_lm, sorted (|m|, l, cos<sin), pruned */
    double *poly;              /* [n_lm][n_int][6] (reference layout)     */
} orc_mp;

/* ---- normalisation constants (reference kernel :54-84, closed form :49-51)
 * PREFACT[m] = sqrt((2m+1) / (4 pi (2m)!)),  COEF[m] = (-1)^m (2m-1)!! PREFACT[m] */
static double g_prefact[ORC_LMAX + 1], g_coef[ORC_LMAX + 1];
static int g_consts_ready = 0;

static void orc_init_consts(void)
{
    if (g_consts_ready) return;
    for (int m = 0; m <= ORC_LMAX; m++) {
        long double lf = 0.0L;                    /* log (2m)! */
        for (int j = 2; j <= 2 * m; j++) lf += logl((long double)j);
        long double pf = sqrtl((2.0L * m + 1.0L) / (4.0L * 3.14159265358979323846264338327950288L)
                               * expl(-lf));
        long double dfac = 1.0L;                  /* (2m-1)!! */
        for (int j = 2 * m - 1; j > 1; j -= 2) dfac *= (long double)j;
        g_prefact[m] = (double)pf;
        g_coef[m] = (double)((m & 1 ? -1.0L : 1.0L) * dfac * pf);
    }
    g_consts_ready = 1;
}

/* Angular factors of the reference kernel (:86-88): 2 sqrt(pi) for m = 0 and
 * 2 sqrt(2 pi) for m != 0.  The reference's literal for the latter,
 * 5.0132565706694072, is 4.3e-9 too large relative to 2 sqrt(2 pi); the exact
 * value is used here. */
#define MUL0 3.5449077018110320546   /* 2 sqrt(pi)   */
#define MUL1 5.0132565492620004855   /* 2 sqrt(2 pi) */

/* ------------------------------------------------------------------------ */
/*  Table build                                                               */
/* ------------------------------------------------------------------------ */

/* Tridiagonal solve (Thomas) of the quintic C2 system, gpu_potential.py:351-401.
 * f, df, d2: [n][k] row-major. */
static void quintic_d2(const double *x, int n, int k, const double *f, const double *df, double *d2)
{
    double *dg = calloc(n, sizeof(double)), *up = calloc(n, sizeof(double)), *lo = calloc(n, sizeof(double));
    double *rhs = calloc((size_t)n * k, sizeof(double));
    for (int i = 0; i < n; i++) {
        if (i > 0) {
            double h = 1.0 / (x[i] - x[i - 1]);
            dg[i] += 3.0 * h; lo[i] = -h;
            for (int c = 0; c < k; c++)
                rhs[i * k + c] -= (20.0 * (f[i * k + c] - f[(i - 1) * k + c]) * h
                                   - 12.0 * df[i * k + c] - 8.0 * df[(i - 1) * k + c]) * h * h;
        }
        if (i < n - 1) {
            double h = 1.0 / (x[i + 1] - x[i]);
            dg[i] += 3.0 * h; up[i] = -h;
            for (int c = 0; c < k; c++)
                rhs[i * k + c] += (20.0 * (f[(i + 1) * k + c] - f[i * k + c]) * h
                                   - 12.0 * df[i * k + c] - 8.0 * df[(i + 1) * k + c]) * h * h;
        }
    }
    double h0 = 1.0 / (x[1] - x[0]), hn = 1.0 / (x[n - 1] - x[n - 2]);
    up[0] = -2.0 * h0;
    lo[n - 1] = -2.0 * hn;
    for (int c = 0; c < k; c++) {
        rhs[c] = (30.0 * (f[k + c] - f[c]) * h0 - 14.0 * df[k + c] - 16.0 * df[c]) * h0 * h0;
        rhs[(n - 1) * k + c] = (-30.0 * (f[(n - 1) * k + c] - f[(n - 2) * k + c]) * hn
                                + 14.0 * df[(n - 2) * k + c] + 16.0 * df[(n - 1) * k + c]) * hn * hn;
    }
    /* forward sweep */
    double *cp = calloc(n, sizeof(double));
    cp[0] = up[0] / dg[0];
    for (int c = 0; c < k; c++) rhs[c] /= dg[0];
    for (int i = 1; i < n; i++) {
        double den = dg[i] - lo[i] * cp[i - 1];
        cp[i] = (i < n - 1) ? up[i] / den : 0.0;
        for (int c = 0; c < k; c++)
            rhs[i * k + c] = (rhs[i * k + c] - lo[i] * rhs[(i - 1) * k + c]) / den;
    }
    for (int c = 0; c < k; c++) d2[(n - 1) * k + c] = rhs[(n - 1) * k + c];
    for (int i = n - 2; i >= 0; i--)
        for (int c = 0; c < k; c++)
            d2[i * k + c] = rhs[i * k + c] - cp[i] * d2[(i + 1) * k + c];
    free(dg); free(up); free(lo); free(rhs); free(cp);
}

/* Inner l=0 law, gpu_potential.py:183-263 (see file header for s == v). */
static void inner_law(const double *p0, const double *d0, const double *r, int nR,
                      double *invPhi0, double *s, double *U, double *W, int *logm)
{
    double P1 = p0[0], P2 = p0[1], d1 = d0[0], d2 = d0[1], r1 = r[0], r2 = r[1];
    double lnr = log(r2 / r1), num1 = r1 * d1, safety = 100.0 * 2.220446049250313e-16;
    *invPhi0 = 0.0; *s = 0.0; *U = 0.0; *W = P1; *logm = 0;
    if (num1 == 0.0 || fabs(P1 - P2) < safety * fmax(fabs(P1), fabs(P2))) return;
    double A = lnr * num1 / (P1 - P2);
    if (!isfinite(A) || A >= 0.0) return;
    double ss = 0.0, UU = num1, WW = P1; int lg = 1;
    double pred_pow = num1 / r2;
    double pred_cub = (r2 / r1 * (6.0 * r1 * (P2 - P1) / (r2 - r1) - d1 * (2 * r1 + r2))) / (2 * r2 + r1);
    if (fabs(d2 - pred_cub) < fabs(d2 - pred_pow)) { ss = 2.0; UU = 0.5 * num1; WW = P1 - UU; lg = 0; }
    *s = ss; *U = UU; *W = WW; *logm = lg;
    if (ss <= 0.0) return;
    double inv = 1.0 / WW;
    for (int i = 0; i < nR; i++) if (p0[i] * inv >= 1.0) { inv = 0.0; break; }
    *invPhi0 = inv;
}

/* Outer l=0 law, gpu_potential.py:266-328. */
static void outer_law(const double *p0, const double *d0, const double *r, int nR,
                      double *s, double *U, double *W, int *logm)
{
    double v = -1.0, P1 = p0[nR - 1], P2 = p0[nR - 2], d1 = d0[nR - 1];
    double r1 = r[nR - 1], r2 = r[nR - 2], lnr = log(r2 / r1);
    double num1 = r1 * d1, num2 = v * P1, den1 = P1, den2 = P2 * exp(-v * lnr);
    double safety = 100.0 * 2.220446049250313e-16;
    int roundoff = fabs(num1 - num2) < fmax(fabs(num1), fabs(num2)) * safety ||
                   fabs(den1 - den2) < fmax(fabs(den1), fabs(den2)) * safety;
    *s = 0.0; *U = 0.0; *W = P1; *logm = 0;
    if (!isfinite(P1) || roundoff || den1 == den2) return;
    double A = lnr * (num1 - num2) / (den1 - den2);
    if (!isfinite(A) || A >= 0.0) return;
    *s = v; *U = num1 - v * P1; *W = P1; *logm = 1;
}

static int lm_cmp_key(int l, int m) { int am = m < 0 ? -m : m; return (am * 64 + l) * 2 + (m < 0); }

/* Build the table from raw (r, Phi, dPhi/dr), arrays [nR][ncol] row-major.
 * Returns 0 on success, negative on invalid input. */
int orc_mp_build(const double *r_in, const double *phi_in, const double *dphi_in, int nR, int ncol,
                 const int *l_of, const int *m_of, double prune, orc_mp *T)
{
    orc_init_consts();
    memset(T, 0, sizeof(*T));
    if (nR < 2 || ncol < 1) return -1;
    int nr = nR;
    double *r = malloc(sizeof(double) * nr), *lr = malloc(sizeof(double) * nr);
    double *ph = malloc(sizeof(double) * (size_t)nr * ncol), *dl = malloc(sizeof(double) * (size_t)nr * ncol);
    for (int i = 0; i < nr; i++) { r[i] = r_in[i]; lr[i] = log(r[i]); }
    for (int i = 0; i < nr; i++)
        for (int c = 0; c < ncol; c++) {
            ph[i * ncol + c] = phi_in[i * ncol + c];
            dl[i * ncol + c] = dphi_in[i * ncol + c] * r[i];           /* dPhi/dln r */
        }
    double mean = (lr[nr - 1] - lr[0]) / (nr - 1), spread = 0.0;
    {
        double acc = 0.0;
        for (int i = 0; i + 1 < nr; i++) acc += lr[i + 1] - lr[i];
        mean = acc / (nr - 1);
        for (int i = 0; i + 1 < nr; i++) spread = fmax(spread, fabs((lr[i + 1] - lr[i]) - mean));
        spread /= mean;
    }
    double *d0dr = malloc(sizeof(double) * nr);
    for (int i = 0; i < nr; i++) d0dr[i] = dphi_in[i * ncol + 0];
    if (spread > 1e-3) {
        /* cubic Hermite resample onto >= 1000 log-uniform nodes (gpu_potential.py:468-489) */
        int nn = nr > 1000 ? nr : 1000;
        double *lr2 = malloc(sizeof(double) * nn), *r2 = malloc(sizeof(double) * nn);
        double *ph2 = malloc(sizeof(double) * (size_t)nn * ncol), *dl2 = malloc(sizeof(double) * (size_t)nn * ncol);
        for (int j = 0; j < nn; j++) {
            double x = lr[0] + (lr[nr - 1] - lr[0]) * j / (double)(nn - 1);
            if (j == nn - 1) x = lr[nr - 1];
            lr2[j] = x; r2[j] = exp(x);
            int i = 0;
            while (i < nr - 2 && x > lr[i + 1]) i++;
            double h = lr[i + 1] - lr[i], t = (x - lr[i]) / h;
            double h00 = (1 + 2 * t) * (1 - t) * (1 - t), h10 = t * (1 - t) * (1 - t);
            double h01 = t * t * (3 - 2 * t), h11 = t * t * (t - 1);
            double g00 = 6 * t * t - 6 * t, g10 = 3 * t * t - 4 * t + 1, g01 = -6 * t * t + 6 * t, g11 = 3 * t * t - 2 * t;
            for (int c = 0; c < ncol; c++) {
                double f0 = ph[i * ncol + c], f1 = ph[(i + 1) * ncol + c];
                double m0 = dl[i * ncol + c], m1 = dl[(i + 1) * ncol + c];
                ph2[j * ncol + c] = h00 * f0 + h10 * h * m0 + h01 * f1 + h11 * h * m1;
                dl2[j * ncol + c] = (g00 * f0 + g10 * h * m0 + g01 * f1 + g11 * h * m1) / h;
            }
        }
        free(r); free(lr); free(ph); free(dl); free(d0dr);
        r = r2; lr = lr2; ph = ph2; dl = dl2; nr = nn;
        d0dr = malloc(sizeof(double) * nr);
        for (int i = 0; i < nr; i++) d0dr[i] = dl[i * ncol] / r[i];
        double acc = 0.0;
        for (int i = 0; i + 1 < nr; i++) acc += lr[i + 1] - lr[i];
        mean = acc / (nr - 1);
    }
    T->logr_min = lr[0];
    T->dlogr = mean;
    T->inv_dlogr = 1.0 / mean;
    T->n_int = nr - 1;
    {
        double *p0 = malloc(sizeof(double) * nr);
        for (int i = 0; i < nr; i++) p0[i] = ph[i * ncol];
        outer_law(p0, d0dr, r, nr, &T->outer_s, &T->outer_U, &T->outer_W, &T->outer_log);
        free(p0);
    }
    /* order columns by (|m|, l, cos<sin) -- gpu_potential.py:513-520 */
    int *ord = malloc(sizeof(int) * ncol);
    for (int c = 0; c < ncol; c++) ord[c] = c;
    for (int a = 1; a < ncol; a++) {                 /* stable insertion sort */
        int v = ord[a], kv = lm_cmp_key(l_of[v], m_of[v]), b = a - 1;
        while (b >= 0 && lm_cmp_key(l_of[ord[b]], m_of[ord[b]]) > kv) { ord[b + 1] = ord[b]; b--; }
        ord[b + 1] = v;
    }
    int lmax_all = 0;
    for (int c = 0; c < ncol; c++) if (l_of[c] > lmax_all) lmax_all = l_of[c];
    if (lmax_all > ORC_LMAX) goto fail;
    if (!(l_of[ord[0]] == 0 && m_of[ord[0]] == 0)) goto fail;
    /* prune columns whose |Phi| < threshold on every node, keep (0,0) -- :528-538 */
    int *keep = malloc(sizeof(int) * ncol), nk = 0;
    for (int a = 0; a < ncol; a++) {
        int c = ord[a], ok = (a == 0) || prune <= 0.0;
        for (int i = 0; i < nr && !ok; i++) if (fabs(ph[i * ncol + c]) >= prune) ok = 1;
        if (ok) keep[nk++] = c;
    }
    T->n_lm = nk;
    T->lm_l = malloc(sizeof(int) * nk);
    T->lm_m = malloc(sizeof(int) * nk);
    T->lmax = 0;
    for (int a = 0; a < nk; a++) {
        T->lm_l[a] = l_of[keep[a]]; T->lm_m[a] = m_of[keep[a]];
        if (T->lm_l[a] > T->lmax) T->lmax = T->lm_l[a];
    }
    /* log-scaling -- :543-569 */
    double *f = malloc(sizeof(double) * (size_t)nr * nk), *df = malloc(sizeof(double) * (size_t)nr * nk);
    int ls = 1;
    for (int i = 0; i < nr; i++) if (!(ph[i * ncol + keep[0]] < 0.0)) ls = 0;
    T->log_scaling = ls;
    T->invPhi0 = 0.0; T->inner_s = 0.0; T->inner_U = 0.0; T->inner_W = ph[keep[0]]; T->inner_log = 0;
    if (ls) {
        double *p0 = malloc(sizeof(double) * nr);
        for (int i = 0; i < nr; i++) p0[i] = ph[i * ncol + keep[0]];
        inner_law(p0, d0dr, r, nr, &T->invPhi0, &T->inner_s, &T->inner_U, &T->inner_W, &T->inner_log);
        free(p0);
    }
    for (int i = 0; i < nr; i++) {
        double P0 = ph[i * ncol + keep[0]], D0 = dl[i * ncol + keep[0]];
        for (int a = 0; a < nk; a++) {
            double P = ph[i * ncol + keep[a]], D = dl[i * ncol + keep[a]];
            if (!ls) { f[i * nk + a] = P; df[i * nk + a] = D; }
            else if (a == 0) {
                f[i * nk] = log(T->invPhi0 - 1.0 / P0);
                df[i * nk] = D0 / (P0 * (T->invPhi0 * P0 - 1.0));
            } else {
                double ratio = P / P0;
                f[i * nk + a] = ratio;
                df[i * nk + a] = (D - ratio * D0) / P0;
            }
        }
    }
    double *d2 = malloc(sizeof(double) * (size_t)nr * nk);
    quintic_d2(lr, nr, nk, f, df, d2);
    /* Horner coefficients per (c, k) -- :584-605 */
    double h = T->dlogr;
    int ni = T->n_int;
    T->poly = malloc(sizeof(double) * (size_t)nk * ni * 6);
    for (int a = 0; a < nk; a++)
        for (int k = 0; k < ni; k++) {
            double F0 = f[k * nk + a], F1 = f[(k + 1) * nk + a];
            double M0 = df[k * nk + a] * h, M1 = df[(k + 1) * nk + a] * h;
            double Q0 = d2[k * nk + a] * h * h, Q1 = d2[(k + 1) * nk + a] * h * h;
            double D = F1 - F0, *p = T->poly + ((size_t)a * ni + k) * 6;
            p[0] = F0;
            p[1] = M0;
            p[2] = 0.5 * Q0;
            p[3] = 10.0 * D - 6.0 * M0 - 4.0 * M1 - 1.5 * Q0 + 0.5 * Q1;
            p[4] = -15.0 * D + 8.0 * M0 + 7.0 * M1 + 1.5 * Q0 - Q1;
            p[5] = 6.0 * D - 3.0 * M0 - 3.0 * M1 + 0.5 * (Q1 - Q0);
        }
    free(f); free(df); free(d2); free(keep); free(ord);
    free(r); free(lr); free(ph); free(dl); free(d0dr);
    return 0;
fail:
    free(ord); free(r); free(lr); free(ph); free(dl); free(d0dr);
    return -2;
}

void orc_mp_free(orc_mp *T)
{
    free(T->lm_l); free(T->lm_m); free(T->poly);
    memset(T, 0, sizeof(*T));
}

/* ------------------------------------------------------------------------ */
/*  Evaluation                                                                */
/* ------------------------------------------------------------------------ */

static inline void quintic(const double *poly, int c, int k, int ni, double s,
                           double *v, double *d, double *d2)
{
    const double *a = poly + ((size_t)c * ni + k) * 6;
    *v = a[0] + s * (a[1] + s * (a[2] + s * (a[3] + s * (a[4] + s * a[5]))));
    if (d)  *d  = a[1] + s * (2 * a[2] + s * (3 * a[3] + s * (4 * a[4] + s * 5 * a[5])));
    if (d2) *d2 = 2 * a[2] + s * (6 * a[3] + s * (12 * a[4] + s * 20 * a[5]));
}

/* l=0 extrapolation law with value and first two ln r derivatives.
 * Phi = (W + U t) e^{v t} in log mode, W e^{v t} + U e^{s t} otherwise. */
static inline void law(double v, double s, double U, double W, int logm, double t,
                       double *P, double *D, double *D2)
{
    double ev = (v == 0.0) ? 1.0 : exp(v * t);
    if (logm) {
        double q = W + U * t;
        *P = q * ev;
        *D = (U + v * q) * ev;
        *D2 = (2.0 * v * U + v * v * q) * ev;
    } else {
        double es = (s != 0.0) ? exp(s * t) : 1.0;
        *P = W * ev + U * es;
        *D = v * W * ev + s * U * es;
        *D2 = v * v * W * ev + s * s * U * es;
    }
}

enum { ORC_POT = 0, ORC_FORCE = 1, ORC_HESS = 2, ORC_DENS = 3 };

/* Spherical derivative accumulators of the interior sum. */
typedef struct {
    double P, Pr, Prr, Pt, Ptt, Prt, Pf, Pff, Prf, Ptf, Pfos, lap;
} acc_t;

typedef struct {
    double C0, dC0, d2C0, idl, s, ct, st, cm, sm;
    int k, am, near_pole, log_scaling;
} term_ctx;

/* ACCUM_LM / ACCUM_HESS / ACCUM_DEN of the reference kernel
 * (:308-337, :592-632, :890-906) for one active column ci at degree l. */
static void accum(const orc_mp *T, const term_ctx *X, acc_t *A, int ci, int l,
                  double Pl, double dPl, double pfc)
{
    int am = X->am, m = T->lm_m[ci];
    double Y = Pl * pfc, dY = dPl * pfc;
    double cc, dcc, d2cc;
    quintic(T->poly, ci, X->k, T->n_int, X->s, &cc, &dcc, &d2cc);
    double dl_ = dcc * X->idl, d2l_ = d2cc * X->idl * X->idl;
    double Cv = cc, dv = dl_, d2v = d2l_;
    if (X->log_scaling) {
        d2v = d2l_ * X->C0 + 2.0 * dl_ * X->dC0 + cc * X->d2C0;
        dv = dl_ * X->C0 + cc * X->dC0;
        Cv = cc * X->C0;
    }
    double mul = am == 0 ? MUL0 : MUL1;
    int is_cos = m >= 0;
    double Tm = is_cos ? X->cm : X->sm;
    double dTm = is_cos ? -(double)am * X->sm : (double)am * X->cm;
    double d2Tm = -(double)(am * am) * Tm;
    double d2Y = X->near_pole ? 0.0
        : (-X->ct / X->st * dY - ((double)(l * (l + 1)) - (double)(am * am) / (X->st * X->st)) * Y);
    A->P   += mul * Cv * Y * Tm;
    A->Pr  += mul * dv * Y * Tm;
    A->Prr += mul * d2v * Y * Tm;
    A->Pt  += mul * Cv * dY * Tm;
    A->Ptt += mul * Cv * d2Y * Tm;
    A->Prt += mul * dv * dY * Tm;
    A->Pf  += mul * Cv * Y * dTm;
    A->Pff += mul * Cv * Y * d2Tm;
    A->Prf += mul * dv * Y * dTm;
    A->Ptf += mul * Cv * dY * dTm;
    /* pole limit of Y / sin(theta): +dY/dtheta at the north pole, -dY/dtheta at
     * the south pole (the reference takes +dY/dtheta at both, kernel :332-334) */
    double Pos = (am == 0) ? 0.0 : (!X->near_pole ? Y / X->st
                                    : (am == 1 ? (X->ct < 0.0 ? -dY : dY) : 0.0));
    A->Pfos += mul * Cv * Pos * dTm;
    A->lap += mul * (d2v + dv - (double)(l * (l + 1)) * Cv) * Y * Tm;
}

static inline int iabs_(int v) { return v < 0 ? -v : v; }

/* The Hessian on the z axis (sin(theta) < 1e-10), where the spherical chain
 * rule below is singular and the reference's pole branch (d2Y = 0, kernel
 * :610-620, and no |m| = 1, 2 limits) gives a wrong H_xx, H_yy, H_xy, H_xz and
 * H_yz.  Only |m| <= 2 have second derivatives on the axis.  With
 * Qh_lm = P~_lm / sin^m(theta) at u = cos(theta) = s = +-1 in closed form
 * (Condon-Shortley phase, N_l = sqrt((2l+1)/4pi)):
 *   m = 0: Qh = N_l s^l
 *   m = 1: Qh = -N_l sqrt(l(l+1)) / 2 * s^(l+1)        (-N_l1 P_l'(s))
 *   m = 2: Qh = N_l sqrt((l-1)l(l+1)(l+2)) / 8 * s^l    (N_l2 P_l''(s))
 * a term is F(r) mul_m Qh(z/r) h_m(x, y) / r^m with h_0 = 1, h_1 = x (cos)
 * or y (sin), h_2 = x^2 - y^2 (cos) or 2xy (sin); differentiating at x = y = 0
 * with V = mul F Qh and t = ln r:
 *   m = 0: H_zz = (V_tt - V_t)/r^2, H_xx = H_yy = (V_t - l(l+1) V / 2)/r^2
 *   m = 1: H_xz (cos), H_yz (sin) = s (V_t - V)/r^2
 *   m = 2: H_xx = -H_yy = 2 V / r^2 (cos), H_xy = 2 V / r^2 (sin). */
static void pole_hessian(const orc_mp *T, const term_ctx *X, double r2, double *out_h)
{
    const double s = X->ct < 0.0 ? -1.0 : 1.0;
    const double pf0 = MUL0 * g_prefact[0];
    double v0t = pf0 * X->dC0, v0tt = pf0 * X->d2C0, l0 = 0.0;
    double v1[2] = {0.0, 0.0}, r1[2] = {0.0, 0.0}, v2[2] = {0.0, 0.0};   /* [cos, sin] */
    for (int ci = 1; ci < T->n_lm; ci++) {
        const int l = T->lm_l[ci], m = T->lm_m[ci], am = iabs_(m);
        if (am > 2) continue;
        double cc, dcc, d2cc;
        quintic(T->poly, ci, X->k, T->n_int, X->s, &cc, &dcc, &d2cc);
        double dl_ = dcc * X->idl, d2l_ = d2cc * X->idl * X->idl;
        double Cv = cc, dv = dl_, d2v = d2l_;
        if (X->log_scaling) {
            d2v = d2l_ * X->C0 + 2.0 * dl_ * X->dC0 + cc * X->d2C0;
            dv = dl_ * X->C0 + cc * X->dC0;
            Cv = cc * X->C0;
        }
        const double Nl = sqrt((2.0 * l + 1.0) / (4.0 * 3.14159265358979323846)), sl = (l % 2) ? s : 1.0;
        double Qh;
        if (am == 0) Qh = MUL0 * Nl * sl;
        else if (am == 1) Qh = -MUL1 * Nl * sqrt((double)l * (l + 1)) * 0.5 * sl * s;
        else Qh = MUL1 * Nl * sqrt((double)(l - 1) * l * (l + 1) * (l + 2)) * 0.125 * sl;
        const int is_sin = m < 0;
        if (am == 0) {
            v0t += Qh * dv; v0tt += Qh * d2v; l0 += (double)(l * (l + 1)) * Qh * Cv;
        } else if (am == 1) {
            v1[is_sin] += Qh * Cv; r1[is_sin] += Qh * dv;
        } else {
            v2[is_sin] += Qh * Cv;
        }
    }
    const double ir2 = 1.0 / r2, hxx0 = (v0t - 0.5 * l0) * ir2;
    out_h[0] = hxx0 + 2.0 * v2[0] * ir2;
    out_h[1] = hxx0 - 2.0 * v2[0] * ir2;
    out_h[2] = (v0tt - v0t) * ir2;
    out_h[3] = 2.0 * v2[1] * ir2;
    out_h[4] = s * (r1[1] - v1[1]) * ir2;
    out_h[5] = s * (r1[0] - v1[0]) * ir2;
}

/* One particle.  out_g: gradient (+grad Phi), out_h: [Hxx,Hyy,Hzz,Hxy,Hyz,Hxz] */
static void eval_one(const orc_mp *T, double px, double py, double pz, int mode,
                     double *out_p, double *out_g, double *out_h, double inv4piG)
{
    const double pf0 = MUL0 * g_prefact[0];
    double r2 = px * px + py * py + pz * pz, r = sqrt(r2), R = sqrt(px * px + py * py);
    int i;
    if (out_g) for (i = 0; i < 3; i++) out_g[i] = 0.0;
    if (out_h) for (i = 0; i < 6; i++) out_h[i] = 0.0;
    if (r < 1e-300) {                           /* origin: kernel :166-171, :829 */
        *out_p = (mode == ORC_DENS) ? 0.0 : pf0 * T->inner_W;
        return;
    }
    double lr = log(r);
    int outer = 0, inner = lr < T->logr_min;
    int k = 0;
    if (!inner) {
        k = (int)((lr - T->logr_min) * T->inv_dlogr);
        if (k < 0) k = 0;
        if (k >= T->n_int) outer = 1;
    }
    if (inner || outer) {                       /* kernel :175-214, :504-547, :832-851 */
        double P, D, D2;
        if (inner) law(0.0, T->inner_s, T->inner_U, T->inner_W, T->inner_log, lr - T->logr_min, &P, &D, &D2);
        else {
            if (mode == ORC_DENS) { *out_p = 0.0; return; }
            law(-1.0, T->outer_s, T->outer_U, T->outer_W, T->outer_log,
                lr - (T->logr_min + T->n_int * T->dlogr), &P, &D, &D2);
        }
        if (mode == ORC_DENS) { *out_p = pf0 * (D2 + D) / (r * r) * inv4piG; return; }
        *out_p = pf0 * P;
        double ir2 = 1.0 / r2;
        if (mode >= ORC_FORCE) {
            double g = pf0 * D;
            out_g[0] = g * px * ir2; out_g[1] = g * py * ir2; out_g[2] = g * pz * ir2;
        }
        if (mode == ORC_HESS) {
            /* Hessian of a radial function with D = dPhi/dln r, D2 = d2Phi/dln r^2:
             * H_ij = (D2 - 2 D) x_i x_j / r^4 + D delta_ij / r^2.  The reference
             * (:514-521, :542-547) uses (D2 - D), which fails the finite-difference
             * check of its own force; the correct coefficient is used here. */
            double A = pf0 * (D2 - 2.0 * D) * ir2 * ir2, B = pf0 * D * ir2, c[3] = {px, py, pz};
            int ii[6] = {0, 1, 2, 0, 1, 0}, jj[6] = {0, 1, 2, 1, 2, 2};
            for (i = 0; i < 6; i++) out_h[i] = A * c[ii[i]] * c[jj[i]] + (ii[i] == jj[i] ? B : 0.0);
        }
        return;
    }
    term_ctx X;
    X.k = k;
    X.idl = T->inv_dlogr;
    X.s = (lr - (T->logr_min + k * T->dlogr)) * T->inv_dlogr;
    if (X.s < 0.0) X.s = 0.0;
    if (X.s > 1.0) X.s = 1.0;
    X.ct = pz / r; X.st = R / r;
    double ph = atan2(py, px), cph = cos(ph), sph = sin(ph);
    X.near_pole = X.st < 1e-10;
    X.log_scaling = T->log_scaling;

    /* monopole with the log-scaling inverse -- kernel :239-261, :569-587 */
    double c0, dc0, d2c0;
    quintic(T->poly, 0, k, T->n_int, X.s, &c0, &dc0, &d2c0);
    double C0 = c0, dC0 = dc0 * X.idl, d2C0 = d2c0 * X.idl * X.idl;
    if (T->log_scaling) {
        double ex = exp(c0), P0 = 1.0 / (T->invPhi0 - ex), dPdX = P0 * P0 * ex;
        d2C0 = dPdX * (d2C0 + dC0 * dC0 * P0 * (T->invPhi0 + ex));
        dC0 = dPdX * dC0;
        C0 = P0;
    }
    X.C0 = C0; X.dC0 = dC0; X.d2C0 = d2C0;
    acc_t A;
    memset(&A, 0, sizeof(A));
    A.P = pf0 * C0; A.Pr = pf0 * dC0; A.Prr = pf0 * d2C0;
    A.lap = pf0 * (d2C0 + dC0);

    /* active terms, walked as in the reference kernel :273-386 */
    double ct = X.ct, st = X.st;
    double sinpow = 1.0;
    X.cm = 1.0; X.sm = 0.0;
    int ci = 1, n = T->n_lm;
    for (int am = 0; am <= T->lmax && ci < n; am++) {
        X.am = am;
        if (iabs_(T->lm_m[ci]) == am) {
            double pf = g_prefact[am], pfc = pf, rp, dp;
            if (am == 0) { rp = 1.0; dp = 0.0; }
            else if (am == 1) { rp = -st; dp = -ct; }
            else { rp = (pf != 0.0) ? g_coef[am] * sinpow / pf : 0.0; dp = X.near_pole ? 0.0 : am * (ct / st) * rp; }
            double rc = rp * ct * (2 * am + 1);
            double dc = (2 * am + 1) * (ct * dp - st * rp);
            if (T->lm_l[ci] == am) {
                if (T->lm_m[ci] == am) { accum(T, &X, &A, ci, am, rp, dp, pfc); ci++; }
                if (am > 0 && ci < n && T->lm_l[ci] == am && T->lm_m[ci] == -am) { accum(T, &X, &A, ci, am, rp, dp, pfc); ci++; }
            }
            for (int l = am + 1; l <= T->lmax; l++) {
                if (l > am + 1) {
                    double il = 1.0 / (double)(l - am);
                    double nr_ = ((2 * l - 1) * ct * rc - (l + am - 1) * rp) * il;
                    double nd_ = ((2 * l - 1) * (ct * dc - st * rc) - (l + am - 1) * dp) * il;
                    rp = rc; dp = dc; rc = nr_; dc = nd_;
                }
                pfc *= sqrt((2.0 * l + 1) / (2.0 * l - 1) * (double)(l - am) / (double)(l + am));
                if (ci >= n) break;
                if (T->lm_l[ci] != l) { if (iabs_(T->lm_m[ci]) != am) break; continue; }
                if (iabs_(T->lm_m[ci]) != am) break;
                if (T->lm_m[ci] == am) { accum(T, &X, &A, ci, l, rc, dc, pfc); ci++; }
                if (am > 0 && ci < n && T->lm_l[ci] == l && T->lm_m[ci] == -am) { accum(T, &X, &A, ci, l, rc, dc, pfc); ci++; }
                if (ci >= n || iabs_(T->lm_m[ci]) != am) break;
            }
        }
        sinpow = (am == 0) ? st : sinpow * st;
        double t = X.cm * cph - X.sm * sph; X.sm = X.sm * cph + X.cm * sph; X.cm = t;
    }
    if (mode == ORC_DENS) { *out_p = inv4piG * A.lap / r2; return; }
    *out_p = A.P;
    if (mode == ORC_POT) return;
    double ir = 1.0 / r, ir2 = ir * ir;
    double cphi = R > 1e-300 ? px / R : 1.0, sphi = R > 1e-300 ? py / R : 0.0;
    out_g[0] = A.Pr * (px * ir2) + A.Pt * (ct * cphi * ir) + A.Pfos * (-sphi * ir);
    out_g[1] = A.Pr * (py * ir2) + A.Pt * (ct * sphi * ir) + A.Pfos * (cphi * ir);
    out_g[2] = A.Pr * (pz * ir2) + A.Pt * (-st * ir);
    if (mode != ORC_HESS) return;
    if (X.near_pole) { pole_hessian(T, &X, r2, out_h); return; }
    /* full Cartesian chain rule -- reference :728-799 */
    double iR = R > 1e-300 ? 1.0 / R : 0.0;
    double dlr[3] = {px * ir2, py * ir2, pz * ir2};
    double dth[3] = {ct * cphi * ir, ct * sphi * ir, -st * ir};
    double dph[3] = {-sphi * iR, cphi * iR, 0.0};
    double ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir;
    double c3[3] = {px, py, pz};
    double d2lr[3][3], d2f[3][3], df_[3] = {-pz * px * ir3, -pz * py * ir3, ir - pz * pz * ir3};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            d2lr[a][b] = (a == b ? ir2 : 0.0) - 2.0 * c3[a] * c3[b] * ir4;
            /* second derivatives of cos(theta) = z/r */
            d2f[a][b] = -((a == b ? pz : 0.0) + (a == 2 ? c3[b] : 0.0) + (b == 2 ? c3[a] : 0.0)) * ir3
                        + 3.0 * pz * c3[a] * c3[b] * ir5;
        }
    double isth = st > 1e-14 ? 1.0 / st : 0.0, cos3 = ct * isth * isth * isth;
    double d2ph[3][3] = {{0}};
    if (R > 1e-14) {
        double iR4 = iR * iR * iR * iR;
        d2ph[0][0] = 2.0 * px * py * iR4;
        d2ph[0][1] = d2ph[1][0] = (py * py - px * px) * iR4;
        d2ph[1][1] = -2.0 * px * py * iR4;
    }
    int ii[6] = {0, 1, 2, 0, 1, 0}, jj[6] = {0, 1, 2, 1, 2, 2};
    for (int p = 0; p < 6; p++) {
        int a = ii[p], b = jj[p];
        double d2th = -d2f[a][b] * isth - df_[a] * df_[b] * cos3;
        out_h[p] = A.Prr * dlr[a] * dlr[b]
                 + A.Prt * (dlr[a] * dth[b] + dlr[b] * dth[a])
                 + A.Prf * (dlr[a] * dph[b] + dlr[b] * dph[a])
                 + A.Ptt * dth[a] * dth[b]
                 + A.Ptf * (dth[a] * dph[b] + dth[b] * dph[a])
                 + A.Pff * dph[a] * dph[b]
                 + A.Pr * d2lr[a][b] + A.Pt * d2th + A.Pf * d2ph[a][b];
    }
}

/* Evaluate N particles (xyz AoS [N][3]).  mode: 0 pot, 1 pot+grad, 2 pot+grad+hess,
 * 3 density (into phi).  grad/hess may be NULL when not requested.  Outputs are
 * the reference kernels' raw quantities: +grad Phi and +Hessian. */
void orc_mp_eval(const orc_mp *T, const double *xyz, int64_t N, int mode,
                 double *phi, double *grad, double *hess, double inv4piG, int nthreads)
{
    orc_init_consts();
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < N; i++) {
        double g[3], h[6];
        eval_one(T, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], mode, phi + i,
                 (mode == ORC_FORCE || mode == ORC_HESS) ? g : NULL, mode == ORC_HESS ? h : NULL, inv4piG);
        if (grad && (mode == ORC_FORCE || mode == ORC_HESS)) { grad[3 * i] = g[0]; grad[3 * i + 1] = g[1]; grad[3 * i + 2] = g[2]; }
        if (hess && mode == ORC_HESS) for (int p = 0; p < 6; p++) hess[6 * i + p] = h[p];
    }
}

/* Accessors for the ctypes wrapper. */
int orc_mp_sizeof(void) { return (int)sizeof(orc_mp); }
double orc_prefact(int m) { orc_init_consts(); return g_prefact[m]; }
double orc_coef(int m) { orc_init_consts(); return g_coef[m]; }
