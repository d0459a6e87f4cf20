/*
 * oracle.c -- CPU restatement of the reference's batched C2C FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker and never as the thing
 * being measured.
 *
 * Contents (each function cites the reference code it restates):
 *   orc_fft_many        FP64 batched 1D/2D/3D C2C FFT (the role FFTW plays in
 *                       benchmark/accuracy_1d.cpp:12-24 and
 *                       benchmark/accuracy_2d.cpp:14-15): forward = exp(-2 pi i kn/N),
 *                       unnormalised, row-major dims, last dim contiguous.
 *   orc_dft_naive       long-double O(N^2) DFT, the brute-force pin for the above.
 *   orc_generate_data   source/mcfft_1d_utils.cpp:88-100 (glibc rand(), float expr).
 *   orc_gen_rev         source/mcfft_1d_utils.cpp:8-35 (digit-reversal permutation).
 *   orc_ref_metric      benchmark/accuracy_1d.cpp:26-56 (mean clipped relative error).
 *   orc_rel_l2 / orc_max_abs   the error norms the build reports.
 *   orc_ref_fp16_exec   CPU emulation of the reference kernels' arithmetic
 *                       (source/mcfft_1d_device.cpp:10-593) instantiated at
 *                       MCTYPE=half, for the plans the reference supports at
 *                       N in {256, 512, 1024, 131072}; it sets the plain-FP16
 *                       tolerance ("no worse than the reference algorithm").
 *
 * Parity pins: see tests/test_oracle.py (golden vectors from the reference's
 * captured outputs 3.txt / 2.txt / 1.log, brute-force DFT, closed forms,
 * numpy.fft).  orc_ref_fp16_exec is pinned by its algebra (no rounding: equal
 * to orc_fft_many), by FP16-exact inputs (a delta gives exactly all ones at
 * every rounding point), and by a first-order error model: each class of
 * rounding points, switched on alone, must produce the rel-L2 that its
 * rounded tables (computed independently) or the FP16 rounding variance of
 * its values predicts, and the classes must add in quadrature.  What rocWMMA
 * does with a half accumulator cannot be run here (the reference's FP16
 * build does not compile), so it is bracketed: per mma_sync rounding (the
 * default) and FP32 across a complex_mul's four mma_sync (ORC_R_FUSED); the
 * build's plain FP16 is held to the tighter of the two.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long double ldbl;

/* ------------------------------------------------------------------ */
/* Accurate twiddles: exp(sign * 2 pi i * k / n) with exact reduction.  */
/* ------------------------------------------------------------------ */
static void twiddle_ld(long k, long n, int sign, double *re, double *im)
{
    k %= n;
    if (k < 0) k += n;
    ldbl c, s;
    if ((4 * k) % n == 0) { /* exact values at multiples of pi/2 */
        long q = (4 * k) / n;
        c = (q == 0) ? 1 : (q == 2) ? -1 : 0;
        s = (q == 1) ? 1 : (q == 3) ? -1 : 0;
    } else {
        ldbl a = (ldbl)2.0L * (ldbl)3.14159265358979323846264338327950288L * (ldbl)k / (ldbl)n;
        c = cosl(a);
        s = sinl(a);
    }
    *re = (double)c;
    *im = (double)(sign < 0 ? -s : s);
}

/* ------------------------------------------------------------------ */
/* FP64 Stockham mixed-radix FFT of one contiguous sequence.           */
/* ------------------------------------------------------------------ */
typedef struct {
    long n;
    int sign;
    int nf;
    int fac[64];
    double *tw; /* n entries: exp(sign 2 pi i k / n) */
} fftplan;

static void plan_init(fftplan *p, long n, int sign)
{
    p->n = n;
    p->sign = sign;
    p->nf = 0;
    long m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (long f = 3; f * f <= m; f += 2)
        while (m % f == 0) { p->fac[p->nf++] = (int)f; m /= f; }
    if (m > 1) p->fac[p->nf++] = (int)m;
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (long k = 0; k < n; ++k) twiddle_ld(k, n, sign, &p->tw[2 * k], &p->tw[2 * k + 1]);
}

static void plan_free(fftplan *p) { free(p->tw); }

/* One Stockham pass: radix r, current sub-length ns (product of earlier radices).
 * in[j + q*(n/r)] for q in [0,r) -> out[(j/ns)*ns*r + j%ns + q'*ns]  */
static void stockham_pass(const fftplan *p, int r, long ns, const double *in, double *out, double *scratch)
{
    const long n = p->n, m = n / r;
    for (long j = 0; j < m; ++j) {
        long k = j % ns;
        /* twiddle W_{ns*r}^{q*k} = tw[q*k*(n/(ns*r))] */
        long tstep = n / (ns * r);
        double *v = scratch; /* 2r */
        for (int q = 0; q < r; ++q) {
            double xr = in[2 * (j + q * m)], xi = in[2 * (j + q * m) + 1];
            long e = ((long)q * k % (ns * r)) * tstep;
            double wr = p->tw[2 * e], wi = p->tw[2 * e + 1];
            v[2 * q] = xr * wr - xi * wi;
            v[2 * q + 1] = xr * wi + xi * wr;
        }
        long base = (j / ns) * ns * r + k;
        if (r == 2) {
            out[2 * base] = v[0] + v[2];
            out[2 * base + 1] = v[1] + v[3];
            out[2 * (base + ns)] = v[0] - v[2];
            out[2 * (base + ns) + 1] = v[1] - v[3];
        } else if (r == 4) {
            double a0r = v[0] + v[4], a0i = v[1] + v[5];
            double a1r = v[0] - v[4], a1i = v[1] - v[5];
            double b0r = v[2] + v[6], b0i = v[3] + v[7];
            double b1r = v[2] - v[6], b1i = v[3] - v[7];
            /* multiply b1 by sign*i: forward (sign<0) -> -i */
            double jr, ji;
            if (p->sign < 0) { jr = b1i; ji = -b1r; } else { jr = -b1i; ji = b1r; }
            out[2 * base] = a0r + b0r;
            out[2 * base + 1] = a0i + b0i;
            out[2 * (base + ns)] = a1r + jr;
            out[2 * (base + ns) + 1] = a1i + ji;
            out[2 * (base + 2 * ns)] = a0r - b0r;
            out[2 * (base + 2 * ns) + 1] = a0i - b0i;
            out[2 * (base + 3 * ns)] = a1r - jr;
            out[2 * (base + 3 * ns) + 1] = a1i - ji;
        } else {
            long rstep = n / r;
            for (int q2 = 0; q2 < r; ++q2) {
                double sr = 0, si = 0;
                for (int q = 0; q < r; ++q) {
                    long e = ((long)q * q2 % r) * rstep;
                    double wr = p->tw[2 * e], wi = p->tw[2 * e + 1];
                    sr += v[2 * q] * wr - v[2 * q + 1] * wi;
                    si += v[2 * q] * wi + v[2 * q + 1] * wr;
                }
                out[2 * (base + q2 * ns)] = sr;
                out[2 * (base + q2 * ns) + 1] = si;
            }
        }
    }
}

/* in-place FFT of x (2n doubles) using work (2n doubles) */
static void fft1(const fftplan *p, double *x, double *work, double *scratch)
{
    double *a = x, *b = work;
    long ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        stockham_pass(p, p->fac[f], ns, a, b, scratch);
        ns *= p->fac[f];
        double *t = a; a = b; b = t;
    }
    if (a != x) memcpy(x, a, sizeof(double) * 2 * (size_t)p->n);
}

static int max_radix(const fftplan *p)
{
    int r = 4;
    for (int f = 0; f < p->nf; ++f) if (p->fac[f] > r) r = p->fac[f];
    return r;
}

/*
 * Batched multi-dimensional FFT, in place, FP64 interleaved complex.
 * dims[0..rank-1] row-major (dims[rank-1] contiguous); batch transforms of
 * prod(dims) elements stored back to back; sign -1 forward, +1 inverse;
 * unnormalised (FFTW convention used by benchmark/accuracy_1d.cpp:15).
 */
void orc_fft_many(double *data, int rank, const long *dims, long batch, int sign)
{
    long total = 1;
    for (int d = 0; d < rank; ++d) total *= dims[d];
    for (int d = rank - 1; d >= 0; --d) {
        long n = dims[d];
        if (n <= 1) continue;
        long stride = 1;
        for (int e = d + 1; e < rank; ++e) stride *= dims[e];
        long outer = total / (n * stride);
        fftplan p;
        plan_init(&p, n, sign);
        long lines = batch * outer * stride;
        const long CB = (stride > 1) ? 8 : 1; /* columns gathered per line block */
#pragma omp parallel
        {
            double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n * CB);
            double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
            double *scratch = (double *)malloc(sizeof(double) * 2 * (size_t)max_radix(&p));
#pragma omp for schedule(dynamic, 4)
            for (long lb = 0; lb < (lines + CB - 1) / CB; ++lb) {
                long l0 = lb * CB;
                long cnt = (l0 + CB <= lines) ? CB : lines - l0;
                /* lines with the same (batch*outer) block share a base; columns consecutive */
                for (long c = 0; c < cnt; ++c) {
                    long l = l0 + c;
                    long bo = l / stride, s = l % stride;
                    double *base = data + 2 * (bo * n * stride + s);
                    double *col = buf + 2 * n * c;
                    if (stride == 1) {
                        fft1(&p, base, work, scratch);
                        continue;
                    }
                    for (long i = 0; i < n; ++i) {
                        col[2 * i] = base[2 * i * stride];
                        col[2 * i + 1] = base[2 * i * stride + 1];
                    }
                    fft1(&p, col, work, scratch);
                    for (long i = 0; i < n; ++i) {
                        base[2 * i * stride] = col[2 * i];
                        base[2 * i * stride + 1] = col[2 * i + 1];
                    }
                }
            }
            free(buf);
            free(work);
            free(scratch);
        }
        plan_free(&p);
    }
}

/*
 * One long 1D FFT (n = 2^k, k >= 2) per transform, parallel inside the
 * transform: the textbook four-step factorisation n = n1 * n2,
 *   X[k1 + n1 k2] = sum_n2 W_n2^(n2 k2) W_n^(n2 k1) sum_n1 x[n1 n2_count + n2] W_n1^(n1 k1),
 * with every short FFT done by fft1 above (FP64 Stockham) and the twiddle
 * W_n^(n2 k1) from twiddle_ld.  orc_fft_many runs a single transform on one
 * thread; at 2^24..2^27 points that takes minutes, this minutes / threads.
 * Pinned by equality with orc_fft_many (tests/test_oracle.py).
 */
void orc_fft_1d_large(double *data, long n, long batch, int sign)
{
    int lg = 0;
    while ((1L << lg) < n) ++lg;
    const long n1 = 1L << (lg / 2), n2 = n / n1;
    fftplan p1, p2;
    plan_init(&p1, n1, sign);
    plan_init(&p2, n2, sign);
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (long b = 0; b < batch; ++b) {
        double *x = data + 2 * n * b;
#pragma omp parallel
        {
            double *col = (double *)malloc(sizeof(double) * 2 * (size_t)n1);
            double *work = (double *)malloc(sizeof(double) * 2 * (size_t)(n1 > n2 ? n1 : n2));
            double *scratch = (double *)malloc(sizeof(double) * 2 * 16);
            /* columns n2: length n1 at stride n2, then the twiddle W_n^(k1 n2) */
#pragma omp for schedule(dynamic, 16)
            for (long c = 0; c < n2; ++c) {
                for (long i = 0; i < n1; ++i) {
                    col[2 * i] = x[2 * (i * n2 + c)];
                    col[2 * i + 1] = x[2 * (i * n2 + c) + 1];
                }
                fft1(&p1, col, work, scratch);
                for (long k1 = 0; k1 < n1; ++k1) {
                    double wr, wi;
                    twiddle_ld(k1 * c, n, sign, &wr, &wi);
                    const double re = col[2 * k1], im = col[2 * k1 + 1];
                    x[2 * (k1 * n2 + c)] = re * wr - im * wi;
                    x[2 * (k1 * n2 + c) + 1] = re * wi + im * wr;
                }
            }
            /* rows k1: length n2, contiguous; result k2 goes to k1 + n1 k2 */
#pragma omp for schedule(dynamic, 16)
            for (long k1 = 0; k1 < n1; ++k1) {
                double *row = x + 2 * k1 * n2;
                fft1(&p2, row, work, scratch);
                for (long k2 = 0; k2 < n2; ++k2) {
                    t[2 * (k1 + n1 * k2)] = row[2 * k2];
                    t[2 * (k1 + n1 * k2) + 1] = row[2 * k2 + 1];
                }
            }
            free(col);
            free(work);
            free(scratch);
        }
        memcpy(x, t, sizeof(double) * 2 * (size_t)n);
    }
    free(t);
    plan_free(&p1);
    plan_free(&p2);
}

/* Brute-force long-double DFT of one sequence (the pin for orc_fft_many). */
void orc_dft_naive(const double *in, double *out, long n, int sign)
{
    ldbl *c = (ldbl *)malloc(sizeof(ldbl) * (size_t)n), *s = (ldbl *)malloc(sizeof(ldbl) * (size_t)n);
    const ldbl twopi = 2.0L * 3.14159265358979323846264338327950288L;
    for (long k = 0; k < n; ++k) {
        c[k] = cosl(twopi * (ldbl)k / (ldbl)n);
        s[k] = (sign < 0 ? -1 : 1) * sinl(twopi * (ldbl)k / (ldbl)n);
    }
    for (long k = 0; k < n; ++k) {
        ldbl sr = 0, si = 0;
        for (long j = 0; j < n; ++j) {
            long e = (long)(((unsigned long)j * (unsigned long)k) % (unsigned long)n);
            sr += (ldbl)in[2 * j] * c[e] - (ldbl)in[2 * j + 1] * s[e];
            si += (ldbl)in[2 * j] * s[e] + (ldbl)in[2 * j + 1] * c[e];
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
    free(c);
    free(s);
}

/* source/mcfft_1d_utils.cpp:88-100: srand(seed); re then im = 0.0001f*rand()/RAND_MAX
 * evaluated in float (0.0001f * int -> float, / int -> float), stored to double. */
void orc_generate_data(double *data, long n, long n_batch, int seed)
{
    srand((unsigned)seed);
    for (long i = 0; i < n_batch; ++i)
        for (long j = 0; j < n; ++j) {
            float a = 0.0001f * (float)rand();
            a = a / (float)RAND_MAX;
            data[0 + j * 2 + i * n * 2] = (double)a;
            float b = 0.0001f * (float)rand();
            b = b / (float)RAND_MAX;
            data[1 + j * 2 + i * n * 2] = (double)b;
        }
}

/* source/mcfft_1d_utils.cpp:8-35: digit-reversal for the radix list, applied from
 * the last radix to the first; rev[i] is the source index placed at position i. */
void orc_gen_rev(long n, int *rev, const int *radices, int n_radices)
{
    int *t0 = (int *)malloc(sizeof(int) * (size_t)n), *t1 = (int *)malloc(sizeof(int) * (size_t)n);
    long now = n;
    for (long i = 0; i < n; ++i) t0[i] = (int)i;
    for (int i = n_radices - 1; i >= 0; --i) {
        int r = radices[i];
        for (long j = 0; j < n; j += now)
            for (int k = 0; k < r; ++k)
                for (long l = 0; l < now / r; ++l) t1[j + l + k * (now / r)] = t0[j + l * r + k];
        now /= r;
        int *t = t0; t0 = t1; t1 = t;
    }
    memcpy(rev, t0, sizeof(int) * (size_t)n);
    free(t0);
    free(t1);
}

/* benchmark/accuracy_1d.cpp:26-56: sum over re and im of min(1, |t-s|/|s|), / (n*batch) */
double orc_ref_metric(const double *tested, const double *standard, long n_total)
{
    double err = 0;
    for (long j = 0; j < 2 * n_total; ++j) {
        double r = fabs((tested[j] - standard[j]) / standard[j]);
        err += (r < 1.0) ? r : 1.0; /* std::min(1.0, NaN) returns 1.0 */
    }
    return err / (double)n_total;
}

double orc_rel_l2(const double *a, const double *ref, long n_total)
{
    ldbl num = 0, den = 0;
    for (long j = 0; j < 2 * n_total; ++j) {
        ldbl d = (ldbl)a[j] - (ldbl)ref[j];
        num += d * d;
        den += (ldbl)ref[j] * (ldbl)ref[j];
    }
    return (double)sqrtl(num / (den > 0 ? den : 1));
}

double orc_max_abs(const double *a, const double *ref, long n_total)
{
    double m = 0;
    for (long j = 0; j < 2 * n_total; ++j) {
        double d = fabs(a[j] - ref[j]);
        if (!(d <= m)) m = d; /* propagates NaN */
    }
    return m;
}

/* ================================================================== */
/* Emulation of the reference kernels at MCTYPE=half.                  */
/* ================================================================== */

/* Rounding points of the FP16 emulation (bit flags; 0 = FP64 everywhere, used
 * to check the emulator's algebra):
 *   ORC_R_INPUT  input converted to half (setup, source/mcfft_1d_utils.cpp:50-51)
 *   ORC_R_TABLE  DFT matrix and W_N_K twiddles: cosf/sinf of a float argument,
 *                stored as half (source/mcfft_1d_device.cpp:32-36, 777-785)
 *   ORC_R_MMA    the half accumulator of each mma_sync rounded back to half
 *   ORC_R_VALU   half arithmetic of cmul/cadd/csub, the twiddle recurrence and
 *                the radix-2/4 tails (source/mcfft_1d_device.cpp:39-55, 81-82, ...)
 *   ORC_R_FUSED  with ORC_R_MMA: the four mma_sync of one complex_mul
 *                accumulate in FP32 and round once per output (the tighter of
 *                the two accumulator semantics; without it every mma_sync rounds)
 * The default (round = 1 from the API) is INPUT|TABLE|MMA|VALU. */
#define ORC_R_INPUT 1
#define ORC_R_TABLE 2
#define ORC_R_MMA 4
#define ORC_R_VALU 8
#define ORC_R_FUSED 16
static int g_rflags = ORC_R_INPUT | ORC_R_TABLE | ORC_R_MMA | ORC_R_VALU;
static int g_round = 1; /* rounding of the current rounding point (rh) */

/* round a double to the nearest IEEE binary16 value (RNE, subnormals, overflow to inf) */
static double rh(double x)
{
    if (!g_round) return x;
    if (x != x) return x;
    double a = fabs(x);
    if (a >= 65520.0) return x > 0 ? INFINITY : -INFINITY;
    int e;
    frexp(a, &e);              /* a = f * 2^e, f in [0.5,1) */
    int ulp_e = e - 11;        /* 11 significant bits */
    if (ulp_e < -24) ulp_e = -24; /* subnormal spacing 2^-24 */
    double q = ldexp(a, -ulp_e);
    double r = nearbyint(q);   /* default rounding mode: nearest-even */
    return copysign(ldexp(r, ulp_e), x);
}


typedef struct { double x, y; } hc; /* value held in a half2 register */

/* rounding at one class of rounding points */
static double rp(int flag, double x)
{
    if (!(g_rflags & flag)) return x;
    const int saved = g_round;
    g_round = 1;
    const double r = rh(x);
    g_round = saved;
    return r;
}
#define rv(x) rp(ORC_R_VALU, (x))

static hc hmul(hc a, hc b) { hc t = { rv(rv(a.x * b.x) - rv(a.y * b.y)), rv(rv(a.x * b.y) + rv(a.y * b.x)) }; return t; }
static hc hadd(hc a, hc b) { hc t = { rv(a.x + b.x), rv(a.y + b.y) }; return t; }
static hc hsub(hc a, hc b) { hc t = { rv(a.x - b.x), rv(a.y - b.y) }; return t; }

/* source/mcfft_1d_device.cpp:32-36: {cosf(2*M_PI*K*M/N), -sinf(...)}; the double
 * argument is converted to float for cosf/sinf, the float result stored to half. */
static hc W_N_K(long N, long K, long M)
{
    double arg = 2 * M_PI * (double)K * (double)M / (double)N;
    hc t;
    if (!(g_rflags & ORC_R_TABLE)) {
        twiddle_ld((long)(((unsigned long)K * (unsigned long)M) % (unsigned long)N), N, -1, &t.x, &t.y);
        return t;
    }
    float af = (float)arg;
    t.x = rp(ORC_R_TABLE, (double)cosf(af));
    t.y = rp(ORC_R_TABLE, -(double)sinf(af));
    return t;
}

static double g_Fr[16][16], g_Fi[16][16];

/* source/mcfft_1d_device.cpp:777-785 (cosf/sinf of 2*pi*i*j/16, stored as MCTYPE) */
static void init_F(void)
{
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) {
            if (g_rflags & ORC_R_TABLE) {
                float a = (float)(2 * M_PI * i * j / 16);
                g_Fr[i][j] = rp(ORC_R_TABLE, (double)cosf(a));
                g_Fi[i][j] = rp(ORC_R_TABLE, -(double)sinf(a));
            } else {
                twiddle_ld((long)(i * j % 16), 16, -1, &g_Fr[i][j], &g_Fi[i][j]);
            }
        }
}


/* One rocWMMA mma_sync with a half accumulator (source/mcfft_1d_device.cpp:23-29):
 * FP32 MFMA products summed onto the accumulator in FP32 (products of two halves
 * are exact in FP32), then -- unless the complex_mul keeps FP32 across its four
 * mma_sync (ORC_R_FUSED) -- the accumulator rounded back to half.  A is 16x16,
 * b is one 16-entry column. */
static double mma_el(const double A[16][16], int row, const double *b, double acc, int round_out)
{
    if (!(g_rflags & ORC_R_MMA)) {
        double s = acc;
        for (int k = 0; k < 16; ++k) s += A[row][k] * b[k];
        return s;
    }
    float s = (float)acc;
    for (int k = 0; k < 16; ++k) s += (float)(A[row][k] * b[k]);
    return round_out ? rp(ORC_R_MMA, (double)s) : (double)s;
}

/* complex_mul (source/mcfft_1d_device.cpp:10-30) for one 16-point column:
 * out_r = -(Fi*Xi) ; out_r += Fr*Xr ; out_i = Fr*Xi ; out_i += Fi*Xr */
static void cmma(const hc *in, hc *out)
{
    double br[16], bi[16];
    const int each = !(g_rflags & ORC_R_FUSED);  /* round after every mma_sync */
    for (int k = 0; k < 16; ++k) { br[k] = in[k].x; bi[k] = in[k].y; }
    for (int i = 0; i < 16; ++i) {
        double r = mma_el(g_Fi, i, bi, 0.0, each);
        r = -r;
        r = mma_el(g_Fr, i, br, r, 1);
        double m = mma_el(g_Fr, i, bi, 0.0, each);
        m = mma_el(g_Fi, i, br, m, 1);
        out[i].x = r;
        out[i].y = m;
    }
}

/* twiddle for row `row` of a lane group that starts at 4*(row/4) and steps by one
 * multiplicative update per element (source/mcfft_1d_device.cpp:81-82,119-123):
 * t = W(N, col, 4*(row/4)); repeated t = t * W(N, col, 1). */
static hc tw_chain(long N, long col, int row)
{
    hc t = W_N_K(N, col, 4 * (row / 4));
    if (!(g_rflags & (ORC_R_TABLE | ORC_R_VALU))) return W_N_K(N, col, row);
    hc f = W_N_K(N, col, 1);
    for (int j = 0; j < row % 4; ++j) t = hmul(t, f);
    return t;
}

/* layer_256_0 body for one 256-chunk c (digit-reversed input), result in o
 * (source/mcfft_1d_device.cpp:58-137). o may alias c. */
static void l256_chunk(const hc *c, hc *o)
{
    hc Y[16][16]; /* Y[k1][n] */
    hc col[16], res[16];
    for (int n = 0; n < 16; ++n) {
        for (int k = 0; k < 16; ++k) col[k] = c[16 * n + k];
        cmma(col, res);
        for (int k1 = 0; k1 < 16; ++k1) Y[k1][n] = res[k1];
    }
    /* smem[16n + k1] = Y[k1][n]; read back B2[row][cc] = smem[16 row + cc] = Y[cc][row] */
    hc Z[16][16];
    for (int cc = 0; cc < 16; ++cc) {
        for (int row = 0; row < 16; ++row) col[row] = hmul(Y[cc][row], tw_chain(256, cc, row));
        cmma(col, res);
        for (int k2 = 0; k2 < 16; ++k2) Z[k2][cc] = res[k2];
    }
    for (int row = 0; row < 16; ++row)
        for (int cc = 0; cc < 16; ++cc) o[16 * row + cc] = Z[row][cc];
}

/* layer_512_0 (:139-229) and layer_1024_0 (:231-333) bodies for one chunk */
static void l512_chunk(hc *c)
{
    hc s[512];
    l256_chunk(c, s);
    l256_chunk(c + 256, s + 256);
    for (int u = 0; u < 256; ++u) {
        hc e0 = s[u];
        hc e1 = hmul(s[u + 256], W_N_K(512, u, 1));
        c[u] = hadd(e0, e1);
        c[u + 256] = hsub(e0, e1);
    }
}

static void l1024_chunk(hc *c)
{
    hc s[1024];
    for (int q = 0; q < 4; ++q) l256_chunk(c + 256 * q, s + 256 * q);
    for (int u = 0; u < 256; ++u) {
        const int chain = g_rflags & (ORC_R_TABLE | ORC_R_VALU);
        hc t1 = W_N_K(1024, u, 1);
        hc t2 = chain ? hmul(t1, t1) : W_N_K(1024, u, 2);
        hc t3 = chain ? hmul(t2, t1) : W_N_K(1024, u, 3);
        hc e0 = s[u], e1 = hmul(s[u + 256], t1), e2 = hmul(s[u + 512], t2), e3 = hmul(s[u + 768], t3);
        c[u].x = rv(rv(rv(e0.x + e1.x) + e2.x) + e3.x);
        c[u].y = rv(rv(rv(e0.y + e1.y) + e2.y) + e3.y);
        c[u + 256].x = rv(rv(rv(e0.x + e1.y) - e2.x) - e3.y);
        c[u + 256].y = rv(rv(rv(e0.y - e1.x) - e2.y) + e3.x);
        c[u + 512].x = rv(rv(rv(e0.x - e1.x) + e2.x) - e3.x);
        c[u + 512].y = rv(rv(rv(e0.y - e1.y) + e2.y) - e3.y);
        c[u + 768].x = rv(rv(rv(e0.x - e1.y) - e2.x) + e3.y);
        c[u + 768].y = rv(rv(rv(e0.y + e1.x) - e2.y) - e3.x);
    }
}

/* layer_256_1(step) (:335-460) over one group of 256 rows x `step` columns at base d */
static void l256_1_group(hc *d, long step)
{
    hc *Y = (hc *)malloc(sizeof(hc) * 256);
    hc col[16], res[16];
    for (long C = 0; C < step; ++C) {
        long gc = C;
        for (int s = 0; s < 16; ++s) {
            for (int r = 0; r < 16; ++r) col[r] = hmul(d[(16 * s + r) * step + C], tw_chain(16 * step, gc, r));
            cmma(col, res);
            for (int k = 0; k < 16; ++k) Y[16 * s + k] = res[k];
        }
        for (int k = 0; k < 16; ++k) {
            long g2 = gc + step * k;
            for (int s = 0; s < 16; ++s) col[s] = hmul(Y[16 * s + k], tw_chain(256 * step, g2, s));
            cmma(col, res);
            for (int k2 = 0; k2 < 16; ++k2) d[(16 * k2 + k) * step + C] = res[k2];
        }
    }
    free(Y);
}

/*
 * Emulate the reference exec path (setup's host permutation + mcfftExec) at
 * MCTYPE=half for one of the reference plans (source/mcfft_1d_device.cpp:695-728):
 *   256: radices {16,16}, layer_256_0
 *   512: {16,16,2}, layer_512_0
 *   1024: {16,16,4}, layer_1024_0
 *   131072: {16,16,2,16,16}, layer_512_0 then layer_256_1(step=512)
 * x: natural-order FP64 input (n*batch complex), y: natural-order output.
 * round: 0 keeps every value in FP64 (algebra check), 1 rounds at every
 * rounding point (ORC_R_INPUT|TABLE|MMA|VALU), any other value is a set of
 * ORC_R_* flags << 1 (bit 1 = ORC_R_INPUT, ...).  Returns 0, or -1 if n has no
 * reference plan.
 */
int orc_ref_fp16_exec(const double *x, double *y, long n, long batch, int round)
{
    g_rflags = round == 1 ? (ORC_R_INPUT | ORC_R_TABLE | ORC_R_MMA | ORC_R_VALU) : (round >> 1);
    int rad[5], nr;
    switch (n) {
    case 256: rad[0] = 16; rad[1] = 16; nr = 2; break;
    case 512: rad[0] = 16; rad[1] = 16; rad[2] = 2; nr = 3; break;
    case 1024: rad[0] = 16; rad[1] = 16; rad[2] = 4; nr = 3; break;
    case 131072: rad[0] = 16; rad[1] = 16; rad[2] = 2; rad[3] = 16; rad[4] = 16; nr = 5; break;
    default: return -1;
    }
    g_round = 1;
    init_F();
    int *rev = (int *)malloc(sizeof(int) * (size_t)n);
    orc_gen_rev(n, rev, rad, nr);
    hc *d = (hc *)malloc(sizeof(hc) * (size_t)n);
    for (long b = 0; b < batch; ++b) {
        const double *xb = x + 2 * n * b;
        for (long i = 0; i < n; ++i) {
            d[i].x = rp(ORC_R_INPUT, xb[2 * rev[i]]);
            d[i].y = rp(ORC_R_INPUT, xb[2 * rev[i] + 1]);
        }
        if (n == 256) l256_chunk(d, d);
        else if (n == 512) l512_chunk(d);
        else if (n == 1024) l1024_chunk(d);
        else {
            for (long c = 0; c < n; c += 512) l512_chunk(d + c);
            l256_1_group(d, 512);
        }
        double *yb = y + 2 * n * b;
        for (long i = 0; i < n; ++i) { yb[2 * i] = d[i].x; yb[2 * i + 1] = d[i].y; }
    }
    free(d);
    free(rev);
    g_rflags = ORC_R_INPUT | ORC_R_TABLE | ORC_R_MMA | ORC_R_VALU;
    return 0;
}

/* round every value of a buffer to binary16 (RNE) -- the quantisation the
 * FP16 storage format applies to an FP64 input */
void orc_round_half(double *x, long count)
{
    g_round = 1;
    for (long i = 0; i < count; ++i) x[i] = rh(x[i]);
}

/* threads the OpenMP runtime will use (1 when built without OpenMP) */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
