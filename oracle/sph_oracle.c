/*
 * CPU oracle for the SPH step: a sequential restatement of the reference
 * kernels, used only by the tests and by the benchmark's CPU baseline to
 * check and time the HIP engine.  It is never part of the product path.
 *
 * Compiled twice from this file: REAL=float (the FP32 restatement, same
 * operation order as the kernels, mixed-precision spots kept where the
 * reference's double literals force double arithmetic) and REAL=double (the
 * FP64 oracle the precision report is measured against).
 *
 * State is in particle-index order; neighbours are visited the way the
 * reference does it (27 cells in z, y, x offset order, then cellparticles
 * order inside a cell, kernels/common.cl:49-68,354-371), with the grid built
 * by a sequential counting sort, i.e. ascending particle index inside a cell.
 *
 * Definitions where the reference is implementation-defined (stated once,
 * used identically by the engine):
 *   - pow(h, k) and the kernel constants: 315/(64 pi h^9), -45/(pi h^6),
 *     45/(pi h^6) evaluated in double and rounded once;
 *   - length(v) = sqrt((x*x + y*y) + z*z), IEEE sqrt, no FMA contraction;
 *   - rsqrt(x) = 1/sqrt(x); in the FP32 Jacobi, atan/sin/cos are double
 *     polynomials rounded to float (orc_atan_d, orc_sin_d, orc_cos_d), in
 *     the FP64 one half-angle identities (half_atan_cs_d);
 *   - a Givens rotation whose off-diagonal is exactly 0 is the identity;
 *   - cell coordinate: NaN and negative values map to 0, then clamp;
 *   - the SVK Cauchy stress keeps all nine entries.
 * Pairs with both the reference distance and the current distance beyond h
 * contribute exactly zero in the reference (for finite state) and are
 * skipped.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define MAXP 5
#define MAXC 32

typedef struct {
    unsigned n, rx, ry, rz;
    real lo[3], hi[3];
    real h, mass, dt, restdens, stiffness, viscosity, youngs, poisson;
    real gravity[3];
    unsigned nplanes;
    real planes[MAXP][6];
    unsigned nconstraints;
    real constraints[MAXC][6];
    unsigned body_count;
    int polar_fp64;
    int eos_tait;         /* 1: p = k ((rho/rho0)^7 - 1) instead of k (rho - rho0) */
    int mass_consistent;  /* 1: fluid force times the particle's own mass */
    int svk_szz_zero;     /* 1: SVK sigma_zz stored as 0, what the reference's out-of-bounds read of a
                             float[6] gives on its platform (kernels/solid_svk.cl:48,75) */
    int compensated;      /* Kahan-compensated sums (the engine's precision option): bit 0 density and
                             rest density, bit 1 forces, A_pq / A_qq and the displacement gradient */
} orc_params;

/* s += v with Kahan compensation c, the engine's operation sequence */
static void kahan_add(real *s, real *c, real v) {
    real y = v - *c;
    real t = *s + y;
    *c = (t - *s) - y;
    *s = t;
}
/* s += v, compensated when kahan is set */
static void acc_add(real *s, real *c, real v, int kahan) {
    if (kahan) kahan_add(s, c, v);
    else *s = *s + v;
}

typedef struct {
    real h, h2, poly6, spiky, lap;
} kconst;

static kconst make_kconst(const orc_params *p) {
    kconst k;
    double h = (double)p->h;
    double h6 = h * h * h * h * h * h, h9 = h6 * h * h * h;
    double pi = 3.14159265358979323846;
    k.h = p->h;
    k.h2 = p->h * p->h;
    k.poly6 = (real)(315.0 / (64.0 * pi * h9));
    k.spiky = (real)(-45.0 / (pi * h6));
    k.lap = (real)(45.0 / (pi * h6));
    return k;
}

typedef struct {
    real x, y, z;
} vec;

static vec ld(const real *a, unsigned i) {
    vec v = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    return v;
}
static void st(real *a, unsigned i, vec v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}
static vec vsub(vec a, vec b) { vec r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static vec vadd(vec a, vec b) { vec r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static vec vneg(vec a) { vec r = {-a.x, -a.y, -a.z}; return r; }
static vec vscale(real s, vec a) { vec r = {s * a.x, s * a.y, s * a.z}; return r; }
static vec vmuls(vec a, real s) { vec r = {a.x * s, a.y * s, a.z * s}; return r; }
static vec vdivs(vec a, real s) { vec r = {a.x / s, a.y / s, a.z / s}; return r; }
static real vdot(vec a, vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static real vlen(vec a) {
    real s = a.x * a.x + a.y * a.y + a.z * a.z;
    return sizeof(real) == 4 ? (real)sqrtf((float)s) : (real)sqrt((double)s);
}

/* poly6, kernels/common.cl:78-80 */
static real W(real r, const kconst *k) {
    if (r > k->h) return 0;
    real t = k->h2 - r * r;
    return k->poly6 * (t * t * t);
}
/* spiky gradient, kernels/common.cl:81-87 */
static vec gradW(vec v, const kconst *k) {
    real r = vlen(v);
    if (r > k->h) { vec z = {0, 0, 0}; return z; }
    vec dir = vdivs(v, r);
    real s = (k->h - r) * (k->h - r);
    return vmuls(vmuls(dir, k->spiky), s);
}
/* viscosity Laplacian, kernels/common.cl:88-90 */
static real lapW(real r, const kconst *k) {
    if (r > k->h) return 0;
    return k->lap * (k->h - r);
}

/* ---------------------------------------------------------------- grid */
static unsigned grid_axis(real p, real lo, real hi, unsigned res) {
    real v = (p - lo) / (hi - lo) * (real)res;
    unsigned u;
    if (!(v > 0)) u = 0;
    else if (v >= (real)4294967296.0) u = 0xffffffffu;
    else u = (unsigned)v;
    return u < res - 1 ? u : res - 1;
}

/* find_particle_bins + count + exclusive scan + stable insert
 * (kernels/common.cl:372-529 run sequentially). */
void orc_bins(const orc_params *p, const real *pos, unsigned *gridcell, unsigned *gridcount,
              unsigned *celloffset, unsigned *cellparticles) {
    const unsigned ncells = p->rx * p->ry * p->rz;
    unsigned i, c;
    for (i = 0; i < p->n; ++i) {
        unsigned gx = grid_axis(pos[3 * i], p->lo[0], p->hi[0], p->rx);
        unsigned gy = grid_axis(pos[3 * i + 1], p->lo[1], p->hi[1], p->ry);
        unsigned gz = grid_axis(pos[3 * i + 2], p->lo[2], p->hi[2], p->rz);
        gridcell[i] = gz * p->rx * p->ry + gy * p->rx + gx;
    }
    memset(gridcount, 0, sizeof(unsigned) * ncells);
    for (i = 0; i < p->n; ++i) gridcount[gridcell[i]]++;
    unsigned run = 0;
    for (c = 0; c < ncells; ++c) {
        celloffset[c] = run;
        run += gridcount[c];
    }
    unsigned *cursor = (unsigned *)malloc(sizeof(unsigned) * (ncells ? ncells : 1));
    memcpy(cursor, celloffset, sizeof(unsigned) * ncells);
    for (i = 0; i < p->n; ++i) cellparticles[cursor[gridcell[i]]++] = i;
    free(cursor);
}

/* The cells around `key` in the reference's order; returns the count. */
static int cells_around(const orc_params *p, unsigned key, unsigned *cells) {
    unsigned layer = p->rx * p->ry;
    unsigned zo = key / layer, yo = (key - layer * zo) / p->rx, xo = key - layer * zo - p->rx * yo;
    int n = 0, gx, gy, gz;
    for (gz = -1; gz <= 1; ++gz)
        for (gy = -1; gy <= 1; ++gy)
            for (gx = -1; gx <= 1; ++gx) {
                int zn = (int)zo + gz, yn = (int)yo + gy, xn = (int)xo + gx;
                if (xn < 0 || yn < 0 || zn < 0 || xn >= (int)p->rx || yn >= (int)p->ry || zn >= (int)p->rz) continue;
                cells[n++] = (unsigned)zn * layer + (unsigned)yn * p->rx + (unsigned)xn;
            }
    return n;
}

/* --------------------------------------------------------- mirror planes */
static vec plane_pt(const orc_params *p, unsigned s) { vec v = {p->planes[s][0], p->planes[s][1], p->planes[s][2]}; return v; }
static vec plane_nm(const orc_params *p, unsigned s) { vec v = {p->planes[s][3], p->planes[s][4], p->planes[s][5]}; return v; }

/* The reference's per-combination test (kernels/common.cl:658-670), done
 * literally: every set plane must have i and j within h and j at least 1e-5
 * (a double literal) away. */
static int combo_ok(const orc_params *p, unsigned b, vec xi, vec xj) {
    unsigned s;
    for (s = 0; s < p->nplanes; ++s) {
        if (!((b >> s) & 1u)) continue;
        vec pt = plane_pt(p, s), nm = plane_nm(p, s);
        real jd = (real)fabs((double)vdot(nm, vsub(xj, pt)));
        if ((real)fabs((double)vdot(nm, vsub(xi, pt))) > p->h || jd > p->h || (double)jd < 1.0e-5) return 0;
    }
    return 1;
}
static vec reflect_pt(vec a, vec pt, vec nm) {
    real d = vdot(nm, vsub(a, pt));
    return vsub(a, vmuls(vscale(2, nm), d));
}
static vec reflect_dir(vec a, vec nm) {
    real d = vdot(nm, a);
    return vsub(a, vmuls(vscale(2, nm), d));
}
static vec reflect_combo(const orc_params *p, unsigned b, vec x) {
    unsigned s;
    for (s = 0; s < p->nplanes; ++s)
        if ((b >> s) & 1u) x = reflect_pt(x, plane_pt(p, s), plane_nm(p, s));
    return x;
}

/* ------------------------------------------------------------ density */
/* compute_density (kernels/common.cl:610-690); with count == NULL it is
 * compute_original_density (kernels/solids.cl:99-170). */
static void density_impl(const orc_params *p, const real *pos, const unsigned *gridcell, const unsigned *gridcount,
                         const unsigned *celloffset, const unsigned *cellparticles, real *out, unsigned *count) {
    const kconst k = make_kconst(p);
    const unsigned ncombo = 1u << p->nplanes;
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)p->n; ++i) {
        unsigned cells[27];
        int nc = cells_around(p, gridcell[i], cells), c;
        vec xi = ld(pos, (unsigned)i);
        real rho = 0, comp = 0;
        const int kahan = p->compensated & 1;
        unsigned nn = 0;
        for (c = 0; c < nc; ++c) {
            unsigned jp, off = celloffset[cells[c]];
            for (jp = off; jp < off + gridcount[cells[c]]; ++jp) {
                unsigned j = cellparticles[jp], b;
                vec xj = ld(pos, j);
                real r = vlen(vsub(xj, xi));
                if (!(r > k.h)) {
                    if (kahan) kahan_add(&rho, &comp, W(r, &k));
                    else rho = rho + W(r, &k);
                }
                if (r < k.h) ++nn;
                for (b = 1; b < ncombo; ++b) {
                    if (!combo_ok(p, b, xi, xj)) continue;
                    real rr = vlen(vsub(reflect_combo(p, b, xj), xi));
                    if (rr < k.h) {
                        if (kahan) kahan_add(&rho, &comp, W(rr, &k));
                        else rho = rho + W(rr, &k);
                        ++nn;
                    }
                }
            }
        }
        out[i] = rho * p->mass;
        if (count) count[i] = nn;
    }
}

void orc_density(const orc_params *p, const real *pos, const unsigned *gridcell, const unsigned *gridcount,
                 const unsigned *celloffset, const unsigned *cellparticles, real *density, unsigned *nn) {
    unsigned *tmp = nn ? nn : (unsigned *)malloc(sizeof(unsigned) * (p->n ? p->n : 1));
    density_impl(p, pos, gridcell, gridcount, celloffset, cellparticles, density, tmp);
    if (!nn) free(tmp);
}

void orc_original_density(const orc_params *p, const real *pos, const unsigned *gridcell, const unsigned *gridcount,
                          const unsigned *celloffset, const unsigned *cellparticles, real *density0) {
    density_impl(p, pos, gridcell, gridcount, celloffset, cellparticles, density0, NULL);
}

/* ------------------------------------------------------- fluid forces */
/* walls at +-2 m in double (kernels/common.cl:37-43, kernels/fluids.cl:60) */
/* Equation of state: the reference's linear p = k (rho - rho0)
 * (kernels/fluids.cl:28-29), or Tait with exponent 7 when eos_tait is set. */
static real pressure(const orc_params *p, real rho) {
    if (!p->eos_tait) return (rho - p->restdens) * p->stiffness;
    real r = rho / p->restdens;
    real r2 = r * r, r4 = r2 * r2;
    return p->stiffness * (r4 * r2 * r - (real)1);
}

static real wall(real x, real f) {
    double lo = -2.0, hi = 2.0, xd = (double)x;
    if (xd < lo + (double)0.2f) return (real)((double)f + (1.0 / (-lo + xd) - 5.0));
    if (xd > hi - (double)0.2f) return (real)((double)f - (1.0 / (hi - xd) - 5.0));
    return f;
}

/* compute_forces_fluids (kernels/fluids.cl:1-63) */
void orc_forces_fluids(const orc_params *p, const real *pos, const real *vel, const real *rho,
                       const unsigned *gridcell, const unsigned *gridcount, const unsigned *celloffset,
                       const unsigned *cellparticles, real *force) {
    const kconst k = make_kconst(p);
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)p->n; ++i) {
        unsigned cells[27];
        int nc = cells_around(p, gridcell[i], cells), c;
        vec xi = ld(pos, (unsigned)i), vi = ld(vel, (unsigned)i);
        real pi = pressure(p, rho[i]);
        vec f = {0, 0, 0}, fc = {0, 0, 0};
        const int kahan = (p->compensated & 2) != 0;
        for (c = 0; c < nc; ++c) {
            unsigned jp, off = celloffset[cells[c]];
            for (jp = off; jp < off + gridcount[cells[c]]; ++jp) {
                unsigned j = cellparticles[jp];
                if (j == (unsigned)i) continue;
                vec xj = ld(pos, j);
                vec d = vsub(xj, xi);
                real r = vlen(d);
                if (r > k.h) continue;
                vec rv = vsub(ld(vel, j), vi);
                real pj = pressure(p, rho[j]);
                real din = k.h - r;
                real didj = rho[i] * rho[j];
                real vterm = (p->viscosity * din) * k.lap;
                real pterm = (real)((double)(pi + pj) * 0.5 / (double)r * (double)din * (double)din * (double)k.spiky);
                acc_add(&f.x, &fc.x, (d.x * pterm + rv.x * vterm) / didj, kahan);
                acc_add(&f.y, &fc.y, (d.y * pterm + rv.y * vterm) / didj, kahan);
                acc_add(&f.z, &fc.z, (d.z * pterm + rv.z * vterm) / didj, kahan);
            }
        }
        f = vmuls(f, p->mass);
        if (p->mass_consistent) f = vmuls(f, p->mass);
        f.x = wall(xi.x, f.x);
        f.y = wall(xi.y, f.y);
        f.z = wall(xi.z, f.z);
        st(force, (unsigned)i, f);
    }
}

/* ----------------------------------------------------------- integrator */
/* step_forward (kernels/common.cl:691-722) */
void orc_step_forward(const orc_params *p, const real *force, real *vel, real *pos, real *acc,
                      const unsigned *body_number, const unsigned *body_dynamic) {
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)p->n; ++i) {
        if (body_number && body_dynamic) {
            unsigned b = body_number[i];
            if (b > 0 && b <= p->body_count && !body_dynamic[b - 1]) continue;
        }
        vec g = {p->gravity[0], p->gravity[1], p->gravity[2]};
        vec a = vadd(vdivs(ld(force, (unsigned)i), p->mass), g);
        if (acc) st(acc, (unsigned)i, a);
        vec v = ld(vel, (unsigned)i);
        vec vn = vadd(v, vscale(p->dt, a));
        st(pos, (unsigned)i, vadd(ld(pos, (unsigned)i), vscale(p->dt, vn)));
        st(vel, (unsigned)i, vn);
    }
}

/* apply_plane_constraints (kernels/constraints.cl:1-30) */
void orc_plane_constraints(const orc_params *p, const unsigned *mask, real *pos, real *vel) {
    long i;
    for (i = 0; i < (long)p->n; ++i) {
        unsigned c, m = mask[i];
        if (m == 0) continue;
        vec x = ld(pos, (unsigned)i), v = ld(vel, (unsigned)i);
        for (c = 0; c < p->nconstraints; ++c) {
            if (!((m >> c) & 1u)) continue;
            vec pt = {p->constraints[c][0], p->constraints[c][1], p->constraints[c][2]};
            vec nm = {p->constraints[c][3], p->constraints[c][4], p->constraints[c][5]};
            x = vsub(x, vscale(vdot(nm, vsub(x, pt)), nm));
            v = vsub(v, vscale(vdot(nm, v), nm));
        }
        st(pos, (unsigned)i, x);
        st(vel, (unsigned)i, v);
    }
}

/* ------------------------------------------------------------ 3x3 algebra */
/* symmetric storage [00, 11, 22, 01, 02, 12] (kernels/common.cl:27-35) */
static int sidx(int r, int c) {
    if (r == c) return r;
    if (r + c == 1) return 3;
    if (r + c == 2) return 4;
    return 5;
}
static vec sym_mv(const real *s, vec v) {
    vec r = {s[0] * v.x + s[3] * v.y + s[4] * v.z, s[3] * v.x + s[1] * v.y + s[5] * v.z,
             s[4] * v.x + s[5] * v.y + s[2] * v.z};
    return r;
}
static vec mat_mv(const real *m, vec v) {
    vec r = {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
             m[6] * v.x + m[7] * v.y + m[8] * v.z};
    return r;
}

/* Generic-precision helpers for the Jacobi solver (prec = float or double). */
/* cos and sin of angle = 1/2 atan(e / d) (kernels/common.cl:101-155).  FP32:
 * the reference's library calls on the reference's quotient.  FP64 (the
 * precision upgrade, which the reference does not have): the same angle from
 * r = sqrt(d^2 + e^2) without library trig, cos phi = |d| / r,
 * sin phi = sign(d) e / r, cos(phi/2) = sqrt(1/2 + cos(phi)/2),
 * sin(phi/2) = sin phi / (2 cos(phi/2)); the quotient form when r is not
 * finite.  Both are the engine's exact operation sequences. */
/* The FP32 path's atan, sin and cos.  The reference calls the OpenCL
 * platform's library under -cl-fast-relaxed-math, whose results are the
 * platform's own; the engine and this oracle define them once, as double
 * polynomials rounded to float, written out identically on both sides so
 * that their bits agree (pinned against the C library to a float ulp):
 *   atan: |x| > 1 -> pi/2 - atan(1/|x|); then t > tan(pi/12) ->
 *         pi/6 + atan((t sqrt3 - 1) / (t + sqrt3)); then the Taylor series
 *         t sum_{n<=15} (-t^2)^n / (2n+1) (|t| <= 0.268); odd in x;
 *   sin, cos on |a| <= pi/4 + 1e-6 (the Jacobi half angle): Taylor series
 *         to a^19 and a^20. */
double orc_atan_d(double x) {
    double t = fabs(x), u, p, r;
    int inv = t > 1.0, shift = 0, n;
    if (inv) t = 1.0 / t;
    if (t > 0.2679491924311227) {
        t = (t * 1.7320508075688772 - 1.0) / (t + 1.7320508075688772);
        shift = 1;
    }
    u = t * t;
    p = 1.0 / 31.0;
    p = -p;
    for (n = 14; n >= 0; --n) p = p * u + ((n & 1) ? -1.0 : 1.0) / (double)(2 * n + 1);
    r = t * p;
    if (shift) r = 0.5235987755982989 + r;
    if (inv) r = 1.5707963267948966 - r;
    return copysign(r, x);
}
double orc_sin_d(double a) {
    static const double c[10] = {1.0, -1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0,
                                 -1.0 / 39916800.0, 1.0 / 6227020800.0, -1.0 / 1307674368000.0,
                                 1.0 / 355687428096000.0, -1.0 / 121645100408832000.0};
    double u = a * a, p = c[9];
    int n;
    for (n = 8; n >= 0; --n) p = p * u + c[n];
    return a * p;
}
double orc_cos_d(double a) {
    static const double c[11] = {1.0, -1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0,
                                 -1.0 / 3628800.0, 1.0 / 479001600.0, -1.0 / 87178291200.0,
                                 1.0 / 20922789888000.0, -1.0 / 6402373705728000.0,
                                 1.0 / 2432902008176640000.0};
    double u = a * a, p = c[10];
    int n;
    for (n = 9; n >= 0; --n) p = p * u + c[n];
    return p;
}
static void half_atan_cs_f(float e, float d, float *c, float *s) {
    float angle = 0.5f * (float)orc_atan_d((double)(e / d));
    *c = (float)orc_cos_d((double)angle);
    *s = (float)orc_sin_d((double)angle);
}
static void half_atan_cs_d(double e, double d, double *c, double *s) {
    double r = sqrt(d * d + e * e), cphi, sphi, ch;
    if (r <= 1.7e308) {
        double inv = 1.0 / r;
        cphi = fabs(d) * inv;
        sphi = (signbit(d) ? -e : e) * inv;
    } else {
        double x = e / d;
        if (fabs(x) <= 1.0) {
            double q = 1.0 / sqrt(1.0 + x * x);
            cphi = q;
            sphi = x * q;
        } else {
            double ix = 1.0 / fabs(x);
            double q = 1.0 / sqrt(1.0 + ix * ix);
            cphi = ix * q;
            sphi = copysign(q, x);
        }
    }
    ch = sqrt(0.5 + 0.5 * cphi);
    *c = ch;
    *s = sphi / (ch + ch);
}

#define DEFINE_JACOBI(T, SUF, HALF_ATAN_CS, ABS)                                                       \
    static void givens_##SUF(const T *a, const T *p, int od, T *ao, T *po) {                           \
        static const int XX[3] = {0, 0, 1}, YY[3] = {1, 2, 2}, EE[3] = {2, 1, 0}, EX[3] = {4, 3, 3},    \
                         EY[3] = {5, 5, 4};                                                            \
        int k = od - 3, ixx = XX[k], iyy = YY[k], iee = EE[k], iex = EX[k], iey = EY[k], r;             \
        T c = 1, s = 0;                                                                                \
        if (a[od] != 0) HALF_ATAN_CS((T)2 * a[od], a[ixx] - a[iyy], &c, &s);                           \
        T c2 = c * c, s2 = s * s, cs = c * s;                                                          \
        ao[ixx] = a[ixx] * c2 + a[iyy] * s2 + (T)2 * a[od] * cs;                                       \
        ao[od] = -a[ixx] * cs + a[iyy] * cs + a[od] * (c2 - s2);                                       \
        ao[iyy] = a[ixx] * s2 + a[iyy] * c2 - (T)2 * a[od] * cs;                                       \
        ao[iee] = a[iee];                                                                              \
        ao[iex] = a[iex] * c + a[iey] * s;                                                             \
        ao[iey] = -a[iex] * s + a[iey] * c;                                                            \
        for (r = 0; r < 3; ++r) {                                                                      \
            po[3 * r + ixx] = p[3 * r + ixx] * c + p[3 * r + iyy] * s;                                 \
            po[3 * r + iyy] = -p[3 * r + ixx] * s + p[3 * r + iyy] * c;                                \
            po[3 * r + iee] = p[3 * r + iee];                                                          \
        }                                                                                              \
    }                                                                                                  \
    static int maxoff_##SUF(const T *a) {                                                              \
        T m = 0;                                                                                       \
        int im = 0, q;                                                                                 \
        for (q = 0; q < 3; ++q) {                                                                      \
            T v = ABS(a[3 + q]);                                                                       \
            if (v > m) { m = v; im = q; }                                                              \
        }                                                                                              \
        return im + 3;                                                                                 \
    }                                                                                                  \
    static void jacobi_##SUF(const T *in, T *ev, T *p) {                                               \
        T a[6], b[6], q[9];                                                                            \
        int t, it, od;                                                                                 \
        for (t = 0; t < 6; ++t) a[t] = in[t];                                                          \
        for (t = 0; t < 9; ++t) p[t] = (t % 4 == 0) ? 1 : 0;                                           \
        od = maxoff_##SUF(a);                                                                          \
        for (it = 0; it < 8; it += 2) {                                                                \
            givens_##SUF(a, p, od, b, q);                                                              \
            od = maxoff_##SUF(b);                                                                      \
            givens_##SUF(b, q, od, a, p);                                                              \
            od = maxoff_##SUF(a);                                                                      \
        }                                                                                              \
        ev[0] = a[0]; ev[1] = a[1]; ev[2] = a[2];                                                      \
    }                                                                                                  \
    static void pdpt_##SUF(const T *d, const T *p, T *o) {                                             \
        T d0p00 = d[0] * p[0], d1p01 = d[1] * p[1], d2p02 = d[2] * p[2];                               \
        T d0p10 = d[0] * p[3], d1p11 = d[1] * p[4], d2p12 = d[2] * p[5];                               \
        T d0p20 = d[0] * p[6], d1p21 = d[1] * p[7], d2p22 = d[2] * p[8];                               \
        o[0] = p[0] * d0p00 + p[1] * d1p01 + p[2] * d2p02;                                             \
        o[3] = p[0] * d0p10 + p[1] * d1p11 + p[2] * d2p12;                                             \
        o[4] = p[0] * d0p20 + p[1] * d1p21 + p[2] * d2p22;                                             \
        o[1] = p[3] * d0p10 + p[4] * d1p11 + p[5] * d2p12;                                             \
        o[5] = p[3] * d0p20 + p[4] * d1p21 + p[5] * d2p22;                                             \
        o[2] = p[6] * d0p20 + p[7] * d1p21 + p[8] * d2p22;                                             \
    }                                                                                                  \
    static void mas_##SUF(const T *a, const T *s, T *o) {                                              \
        int r, c;                                                                                      \
        for (r = 0; r < 3; ++r)                                                                        \
            for (c = 0; c < 3; ++c)                                                                    \
                o[3 * r + c] = a[3 * r] * s[sidx(0, c)] + a[3 * r + 1] * s[sidx(1, c)] +               \
                               a[3 * r + 2] * s[sidx(2, c)];                                           \
    }

static float fabsf_(float x) { return fabsf(x); }
DEFINE_JACOBI(float, f, half_atan_cs_f, fabsf_)
DEFINE_JACOBI(double, d, half_atan_cs_d, fabs)

/* R = A (A^T A)^(-1/2)   (getR, kernels/common.cl:306-331) */
static void polar(const real *apq, real *rot, int fp64) {
    int k;
    if (fp64) {
        double a[9], s[6], ev[3], pv[9], tmp[6], r[9];
        for (k = 0; k < 9; ++k) a[k] = (double)apq[k];
        s[0] = a[0] * a[0] + a[3] * a[3] + a[6] * a[6];
        s[1] = a[1] * a[1] + a[4] * a[4] + a[7] * a[7];
        s[2] = a[2] * a[2] + a[5] * a[5] + a[8] * a[8];
        s[3] = a[0] * a[1] + a[3] * a[4] + a[6] * a[7];
        s[4] = a[0] * a[2] + a[3] * a[5] + a[6] * a[8];
        s[5] = a[1] * a[2] + a[4] * a[5] + a[7] * a[8];
        jacobi_d(s, ev, pv);
        for (k = 0; k < 3; ++k) ev[k] = 1.0 / sqrt(ev[k]);
        pdpt_d(ev, pv, tmp);
        mas_d(a, tmp, r);
        for (k = 0; k < 9; ++k) rot[k] = (real)r[k];
    } else {
        float a[9], s[6], ev[3], pv[9], tmp[6], r[9];
        for (k = 0; k < 9; ++k) a[k] = (float)apq[k];
        s[0] = a[0] * a[0] + a[3] * a[3] + a[6] * a[6];
        s[1] = a[1] * a[1] + a[4] * a[4] + a[7] * a[7];
        s[2] = a[2] * a[2] + a[5] * a[5] + a[8] * a[8];
        s[3] = a[0] * a[1] + a[3] * a[4] + a[6] * a[7];
        s[4] = a[0] * a[2] + a[3] * a[5] + a[6] * a[8];
        s[5] = a[1] * a[2] + a[4] * a[5] + a[7] * a[8];
        jacobi_f(s, ev, pv);
        for (k = 0; k < 3; ++k) ev[k] = 1.0f / sqrtf(ev[k]);
        pdpt_f(ev, pv, tmp);
        mas_f(a, tmp, r);
        for (k = 0; k < 9; ++k) rot[k] = (real)r[k];
    }
}

/* F = Apq * inverse(Aqq) by Jacobi (jacobiInvert, kernels/common.cl:296-305) */
static void deformation(const real *apq, const real *aqq, real *def, int fp64) {
    int k;
    if (fp64) {
        double m[6], ev[3], pv[9], inv[6], a[9], o[9];
        for (k = 0; k < 6; ++k) m[k] = (double)aqq[k];
        jacobi_d(m, ev, pv);
        for (k = 0; k < 3; ++k) ev[k] = 1.0 / ev[k];
        pdpt_d(ev, pv, inv);
        for (k = 0; k < 9; ++k) a[k] = (double)apq[k];
        mas_d(a, inv, o);
        for (k = 0; k < 9; ++k) def[k] = (real)o[k];
    } else {
        float m[6], ev[3], pv[9], inv[6], a[9], o[9];
        for (k = 0; k < 6; ++k) m[k] = (float)aqq[k];
        jacobi_f(m, ev, pv);
        for (k = 0; k < 3; ++k) ev[k] = (float)(1.0 / (double)ev[k]);
        pdpt_f(ev, pv, inv);
        for (k = 0; k < 9; ++k) a[k] = (float)apq[k];
        mas_f(a, inv, o);
        for (k = 0; k < 9; ++k) def[k] = (real)o[k];
    }
}

/* Exported for the tests that pin the solver against numpy. */
void orc_polar(const real *apq, real *rot, int fp64) { polar(apq, rot, fp64); }
void orc_jacobi_inverse_apply(const real *apq, const real *aqq, real *def, int fp64) { deformation(apq, aqq, def, fp64); }

/* --------------------------------------------------- corotational solid */
/* The sums below take a compensation array c of the same shape (used when
 * kahan is set). */
static void add_apq(real *apq, real *c, vec pv, vec q, real w, int kahan) {
    vec wp = vscale(w, pv);
    acc_add(&apq[0], &c[0], wp.x * q.x, kahan); acc_add(&apq[1], &c[1], wp.x * q.y, kahan);
    acc_add(&apq[2], &c[2], wp.x * q.z, kahan); acc_add(&apq[3], &c[3], wp.y * q.x, kahan);
    acc_add(&apq[4], &c[4], wp.y * q.y, kahan); acc_add(&apq[5], &c[5], wp.y * q.z, kahan);
    acc_add(&apq[6], &c[6], wp.z * q.x, kahan); acc_add(&apq[7], &c[7], wp.z * q.y, kahan);
    acc_add(&apq[8], &c[8], wp.z * q.z, kahan);
}
static void add_aqq(real *aqq, real *c, vec q, real w, int kahan) {
    vec wq = vscale(w, q);
    acc_add(&aqq[0], &c[0], wq.x * q.x, kahan); acc_add(&aqq[3], &c[3], wq.x * q.y, kahan);
    acc_add(&aqq[4], &c[4], wq.x * q.z, kahan); acc_add(&aqq[1], &c[1], wq.y * q.y, kahan);
    acc_add(&aqq[5], &c[5], wq.y * q.z, kahan); acc_add(&aqq[2], &c[2], wq.z * q.z, kahan);
}
static void add_outer(real *o, real *c, vec a, vec b, int kahan) {
    acc_add(&o[0], &c[0], a.x * b.x, kahan); acc_add(&o[1], &c[1], a.x * b.y, kahan);
    acc_add(&o[2], &c[2], a.x * b.z, kahan); acc_add(&o[3], &c[3], a.y * b.x, kahan);
    acc_add(&o[4], &c[4], a.y * b.y, kahan); acc_add(&o[5], &c[5], a.y * b.z, kahan);
    acc_add(&o[6], &c[6], a.z * b.x, kahan); acc_add(&o[7], &c[7], a.z * b.y, kahan);
    acc_add(&o[8], &c[8], a.z * b.z, kahan);
}
static void vacc_add(vec *s, vec *c, vec v, int kahan) {
    acc_add(&s->x, &c->x, v.x, kahan);
    acc_add(&s->y, &c->y, v.y, kahan);
    acc_add(&s->z, &c->z, v.z, kahan);
}

/* compute_rotations_and_strains (kernels/solid_corotation.cl:28-195) */
void orc_rotations_strains(const orc_params *p, const real *pos, const real *opos, const real *density0,
                           const unsigned *body_number, const unsigned *gridcell, const unsigned *gridcount,
                           const unsigned *celloffset, const unsigned *cellparticles, real *rotation, real *strain,
                           real *apq_out) {
    const kconst k = make_kconst(p);
    const unsigned ncombo = 1u << p->nplanes;
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)p->n; ++i) {
        if (body_number && body_number[i] > 0) continue;
        unsigned cells[27];
        int nc = cells_around(p, gridcell[i], cells), c, t;
        vec xi = ld(pos, (unsigned)i), x0i = ld(opos, (unsigned)i);
        real apq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, capq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int kahan = (p->compensated & 2) != 0;
        for (c = 0; c < nc; ++c) {
            unsigned jp, off = celloffset[cells[c]];
            for (jp = off; jp < off + gridcount[cells[c]]; ++jp) {
                unsigned j = cellparticles[jp], b;
                vec xj = ld(pos, j), x0j = ld(opos, j);
                if (j != (unsigned)i) {
                    vec pv = vsub(xj, xi);
                    real r = vlen(pv);
                    if (!(r > k.h)) add_apq(apq, capq, pv, vsub(x0j, x0i), W(r, &k), kahan);
                }
                for (b = 1; b < ncombo; ++b) {
                    if (!combo_ok(p, b, xi, xj)) continue;
                    vec pv = vsub(reflect_combo(p, b, xj), xi);
                    vec q = vsub(reflect_combo(p, b, x0j), x0i);
                    real d = vlen(pv);
                    if (d < k.h) add_apq(apq, capq, pv, q, W(d, &k), kahan);
                }
            }
        }
        if (apq_out) for (t = 0; t < 9; ++t) apq_out[9 * i + t] = apq[t];
        real *rot = rotation + 9 * i;
        polar(apq, rot, p->polar_fp64);
        real rt[9] = {rot[0], rot[3], rot[6], rot[1], rot[4], rot[7], rot[2], rot[5], rot[8]};
        real du[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cdu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        real d0i = density0[i];
        for (c = 0; c < nc; ++c) {
            unsigned jp, off = celloffset[cells[c]];
            for (jp = off; jp < off + gridcount[cells[c]]; ++jp) {
                unsigned j = cellparticles[jp], b;
                vec xj = ld(pos, j), x0j = ld(opos, j);
                if (j != (unsigned)i) {
                    vec q = vsub(x0j, x0i);
                    if (!(vlen(q) > k.h)) {
                        vec u = vsub(mat_mv(rt, vsub(xj, xi)), q);
                        add_outer(du, cdu, vdivs(vmuls(u, p->mass), d0i), gradW(vneg(q), &k), kahan);
                    }
                }
                for (b = 1; b < ncombo; ++b) {
                    if (!combo_ok(p, b, xi, xj)) continue;
                    vec pv = mat_mv(rt, vsub(reflect_combo(p, b, xj), xi));
                    vec q = vsub(reflect_combo(p, b, x0j), x0i);
                    vec u = vsub(pv, q);
                    if (vlen(pv) < k.h) add_outer(du, cdu, vdivs(vmuls(u, p->mass), d0i), gradW(vneg(q), &k), kahan);
                }
            }
        }
        real *s = strain + 6 * i;
        s[0] = du[0];
        s[1] = du[4];
        s[2] = du[8];
        s[3] = (real)0.5 * (du[1] + du[3]);
        s[4] = (real)0.5 * (du[2] + du[6]);
        s[5] = (real)0.5 * (du[5] + du[7]);
    }
}

/* compute_stresses, corotational (kernels/solid_corotation.cl:9-26,197-208) */
void orc_stresses_corot(const orc_params *p, const real *strain, real *stress) {
    real E = p->youngs, nu = p->poisson;
    real lam = E * nu / ((real)1 + nu) / ((real)1 - (real)2 * nu);
    real mu = E / (real)2 / ((real)1 + nu);
    real d = lam + (real)2 * mu, sh = (real)2 * mu;
    real cm[6][6] = {{d, lam, lam, 0, 0, 0}, {lam, d, lam, 0, 0, 0}, {lam, lam, d, 0, 0, 0},
                     {0, 0, 0, sh, 0, 0},    {0, 0, 0, 0, sh, 0},    {0, 0, 0, 0, 0, sh}};
    long i;
    int r;
    for (i = 0; i < (long)p->n; ++i) {
        const real *e = strain + 6 * i;
        for (r = 0; r < 6; ++r)
            stress[6 * i + r] = cm[r][0] * e[0] + cm[r][1] * e[1] + cm[r][2] * e[2] + cm[r][3] * e[3] +
                                cm[r][4] * e[4] + cm[r][5] * e[5];
    }
}

/* ------------------------------------------------------------ SVK solid */
/* compute_deformations (kernels/solid_svk.cl:82-175) */
void orc_deformations_svk(const orc_params *p, const real *pos, const real *opos, const unsigned *gridcell,
                          const unsigned *gridcount, const unsigned *celloffset, const unsigned *cellparticles,
                          real *def) {
    const kconst k = make_kconst(p);
    const unsigned ncombo = 1u << p->nplanes;
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)p->n; ++i) {
        unsigned cells[27];
        int nc = cells_around(p, gridcell[i], cells), c;
        vec xi = ld(pos, (unsigned)i), x0i = ld(opos, (unsigned)i);
        real apq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, aqq[6] = {0, 0, 0, 0, 0, 0};
        real capq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, caqq[6] = {0, 0, 0, 0, 0, 0};
        const int kahan = (p->compensated & 2) != 0;
        for (c = 0; c < nc; ++c) {
            unsigned jp, off = celloffset[cells[c]];
            for (jp = off; jp < off + gridcount[cells[c]]; ++jp) {
                unsigned j = cellparticles[jp], b;
                vec xj = ld(pos, j), x0j = ld(opos, j);
                if (j != (unsigned)i) {
                    vec pv = vsub(xj, xi);
                    real r = vlen(pv);
                    if (!(r > k.h)) {
                        vec q = vsub(x0j, x0i);
                        real w = W(r, &k);
                        add_apq(apq, capq, pv, q, w, kahan);
                        add_aqq(aqq, caqq, q, w, kahan);
                    }
                }
                for (b = 1; b < ncombo; ++b) {
                    if (!combo_ok(p, b, xi, xj)) continue;
                    vec pv = vsub(reflect_combo(p, b, xj), xi);
                    vec q = vsub(reflect_combo(p, b, x0j), x0i);
                    real d = vlen(pv);
                    if (d < k.h) {
                        real w = W(d, &k);
                        add_apq(apq, capq, pv, q, w, kahan);
                        add_aqq(aqq, caqq, q, w, kahan);
                    }
                }
            }
        }
        deformation(apq, aqq, def + 9 * i, p->polar_fp64);
    }
}

/* compute_stresses, SVK (kernels/solid_svk.cl:1-80), full 3x3 Cauchy */
void orc_stresses_svk(const orc_params *p, const real *def, real *stress) {
    long i;
    int r, c;
    for (i = 0; i < (long)p->n; ++i) {
        const real *F = def + 9 * i;
        real Ft[9] = {F[0], F[3], F[6], F[1], F[4], F[7], F[2], F[5], F[8]}, ftf[9], e[6], svk[6], tmp[9], cy[9];
        for (r = 0; r < 3; ++r)
            for (c = 0; c < 3; ++c)
                ftf[3 * r + c] = Ft[3 * r] * F[c] + Ft[3 * r + 1] * F[3 + c] + Ft[3 * r + 2] * F[6 + c];
        e[0] = (real)(0.5 * ftf[0] - 0.5);
        e[1] = (real)(0.5 * ftf[4] - 0.5);
        e[2] = (real)(0.5 * ftf[8] - 0.5);
        e[3] = (real)(0.5 * ftf[1]);
        e[4] = (real)(0.5 * ftf[2]);
        e[5] = (real)(0.5 * ftf[5]);
        real E = p->youngs, nu = p->poisson;
        real lam = (real)((double)(E * nu) / (1.0 + nu) / (1.0 - 2.0 * nu));
        real mu = (real)(E / 2.0 / (1.0 + nu));
        real first = lam * (e[0] + e[1] + e[2]);
        svk[0] = (real)(first + 2.0 * mu * e[0]);
        svk[1] = (real)(first + 2.0 * mu * e[1]);
        svk[2] = (real)(first + 2.0 * mu * e[2]);
        svk[3] = (real)(2.0 * mu * e[3]);
        svk[4] = (real)(2.0 * mu * e[4]);
        svk[5] = (real)(2.0 * mu * e[5]);
        for (r = 0; r < 3; ++r)
            for (c = 0; c < 3; ++c)
                tmp[3 * r + c] = F[3 * r] * svk[sidx(0, c)] + F[3 * r + 1] * svk[sidx(1, c)] + F[3 * r + 2] * svk[sidx(2, c)];
        for (r = 0; r < 3; ++r)
            for (c = 0; c < 3; ++c)
                cy[3 * r + c] = tmp[3 * r] * Ft[c] + tmp[3 * r + 1] * Ft[3 + c] + tmp[3 * r + 2] * Ft[6 + c];
        real det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                   F[2] * (F[3] * F[7] - F[4] * F[6]);
        real sc = (real)(1.0 / det);
        real *s = stress + 6 * i;
        s[0] = sc * cy[0];
        s[1] = sc * cy[4];
        s[2] = p->svk_szz_zero ? (real)0 : sc * cy[8];
        s[3] = sc * cy[1];
        s[4] = sc * cy[2];
        s[5] = sc * cy[5];
    }
}

/* -------------------------------------------------------- solid forces */
/* compute_forces_solids (kernels/solid_corotation.cl:215-403 with rotation,
 * kernels/solid_svk.cl:182-367 with rotation == NULL) */
void orc_forces_solids(const orc_params *p, const real *pos, const real *vel, const real *opos, const real *stress,
                       const real *rotation, const real *density0, const real *density, const unsigned *body_number,
                       const unsigned *gridcell, const unsigned *gridcount, const unsigned *celloffset,
                       const unsigned *cellparticles, real *force, real *force_elastic, real *force_viscous,
                       real *cutoff_force) {
    const kconst k = make_kconst(p);
    const unsigned ncombo = 1u << p->nplanes;
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)p->n; ++i) {
        unsigned cells[27];
        int nc = cells_around(p, gridcell[i], cells), c;
        vec xi = ld(pos, (unsigned)i), vi = ld(vel, (unsigned)i), x0i = ld(opos, (unsigned)i);
        const real *si = stress + 6 * i;
        const real *ri = rotation ? rotation + 9 * i : NULL;
        vec fe = {0, 0, 0}, fv = {0, 0, 0}, fec = {0, 0, 0}, fvc = {0, 0, 0};
        vec cfe = {0, 0, 0}, cfv = {0, 0, 0}, cfec = {0, 0, 0}, cfvc = {0, 0, 0};
        const int kahan = (p->compensated & 2) != 0;
        for (c = 0; c < nc; ++c) {
            unsigned jp, off = celloffset[cells[c]];
            for (jp = off; jp < off + gridcount[cells[c]]; ++jp) {
                unsigned j = cellparticles[jp], b;
                vec x0j = ld(opos, j), xj = ld(pos, j), vj = ld(vel, j);
                vec X = vsub(x0i, x0j);
                real rx = vlen(X), rcur = vlen(vsub(xj, xi));
                int near_pair = !(rx > k.h) || !(rcur > k.h);
                real vivj = p->mass * p->mass / density0[i] / density0[j];
                const real *sj = stress + 6 * j;
                const real *rj = rotation ? rotation + 9 * j : NULL;
                real mrho = p->mass / density[j];
                if (j != (unsigned)i && near_pair) {
                    vec fji = vscale(-vivj, sym_mv(si, gradW(X, &k)));
                    vec fij = (body_number && body_number[j] > 0) ? vneg(fji)
                                                                  : vscale(-vivj, sym_mv(sj, gradW(vneg(X), &k)));
                    vec e = rotation ? vadd(vneg(mat_mv(ri, fji)), mat_mv(rj, fij)) : vadd(vneg(fji), fij);
                    vec v = vmuls(vscale(mrho, vsub(vj, vi)), lapW(rcur, &k));
                    vacc_add(&fe, &cfe, e, kahan);
                    vacc_add(&fv, &cfv, v, kahan);
                    vacc_add(&fec, &cfec, e, kahan);
                    vacc_add(&fvc, &cfvc, v, kahan);
                }
                for (b = 1; b < ncombo; ++b) {
                    if (!combo_ok(p, b, xi, xj)) continue;
                    vec rp = reflect_combo(p, b, xj), rp0 = reflect_combo(p, b, x0j), rip0 = reflect_combo(p, b, x0i);
                    vec Xm = vsub(x0i, rp0), refx = vsub(rip0, x0j);
                    vec fji = vscale(-vivj, sym_mv(si, gradW(Xm, &k)));
                    vec fij = vscale(-vivj, sym_mv(sj, gradW(vneg(refx), &k)));
                    unsigned s, lowest = 0;
                    if (rotation) fij = mat_mv(rj, fij);
                    for (s = p->nplanes; s > 0; --s)
                        if ((b >> (s - 1)) & 1u) {
                            fij = reflect_dir(fij, plane_nm(p, s - 1));
                            lowest = s - 1;
                        }
                    vec e = rotation ? vadd(vneg(mat_mv(ri, fji)), fij) : vadd(vneg(fji), fij);
                    vacc_add(&fe, &cfe, e, kahan);
                    vec refvel = reflect_dir(vj, plane_nm(p, lowest));
                    vec v = vmuls(vscale(mrho, vsub(refvel, vi)), lapW(vlen(vsub(rp, xi)), &k));
                    vacc_add(&fv, &cfv, v, kahan);
                    if (!(b & 1u)) {
                        vacc_add(&fec, &cfec, e, kahan);
                        vacc_add(&fvc, &cfvc, v, kahan);
                    }
                }
            }
        }
        vec e = vmuls(fe, (real)0.5), v = vmuls(fv, p->viscosity);
        st(force, (unsigned)i, vadd(e, v));
        if (force_elastic) st(force_elastic, (unsigned)i, e);
        if (force_viscous) st(force_viscous, (unsigned)i, v);
        if (cutoff_force) st(cutoff_force, (unsigned)i, vadd(vmuls(fec, (real)0.5), vmuls(fvc, p->viscosity)));
    }
}

/* ---- Scripted rigid bodies and initial-condition helpers ------------------
 * populate_position_cuboid and rotate_particles: kernels/common.cl:533-604;
 * set_body_position and set_body_rotation: kernels/rigidbody.cl:1-120.
 * Matrices are row-major 3x3.  Products follow multiplyMatrices
 * (kernels/common.cl:208-220): o[r][c] = (a[r][0] b[0][c] + a[r][1] b[1][c])
 * + a[r][2] b[2][c], zero terms included. */
static real rcos(real x) { return sizeof(real) == 4 ? (real)cosf((float)x) : (real)cos((double)x); }
static real rsin(real x) { return sizeof(real) == 4 ? (real)sinf((float)x) : (real)sin((double)x); }

static void mm3(const real *a, const real *b, real *o) {
    int r, c;
    for (r = 0; r < 3; ++r)
        for (c = 0; c < 3; ++c) o[3 * r + c] = a[3 * r] * b[c] + a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c];
}
static void axis_x(real a, real *m) {
    real c = rcos(a), s = rsin(a);
    m[0] = 1; m[1] = 0; m[2] = 0;
    m[3] = 0; m[4] = c; m[5] = -s;
    m[6] = 0; m[7] = s; m[8] = c;
}
/* The reference's y factor: sin enters as [0][2] = -s, [2][0] = +s. */
static void axis_y(real a, real *m) {
    real c = rcos(a), s = rsin(a);
    m[0] = c; m[1] = 0; m[2] = -s;
    m[3] = 0; m[4] = 1; m[5] = 0;
    m[6] = s; m[7] = 0; m[8] = c;
}
static void axis_z(real a, real *m) {
    real c = rcos(a), s = rsin(a);
    m[0] = c; m[1] = -s; m[2] = 0;
    m[3] = s; m[4] = c; m[5] = 0;
    m[6] = 0; m[7] = 0; m[8] = 1;
}

/* rotate_particles: x <- Rx (Ry Rz) x for the first n particles
 * (kernels/common.cl:561-604). */
void orc_rotate_particles(real *pos, unsigned n, real ax, real ay, real az) {
    real mx[9], my[9], mz[9], myz[9], m[9];
    unsigned i;
    axis_x(ax, mx);
    axis_y(ay, my);
    axis_z(az, mz);
    mm3(my, mz, myz);
    mm3(mx, myz, m);
    for (i = 0; i < n; ++i) st(pos, i, mat_mv(m, ld(pos, i)));
}

/* Body matrix of Euler angles (x, y, x2): Rx(x2) (Ry(y) Rx(x))
 * (get_matrix_from_euler, kernels/rigidbody.cl:35-66). */
void orc_euler_matrix(const real *e, real *out) {
    real m1[9], m2[9], m3[9], t[9];
    axis_x(e[0], m1);
    axis_y(e[1], m2);
    axis_x(e[2], m3);
    mm3(m2, m1, t);
    mm3(m3, t, out);
}

/* populate_position_cuboid: one lane per lattice point (i, j, k), particle
 * index i + sx j + sx sy k in 32-bit unsigned arithmetic; lanes at or past
 * pnum write nothing; the lane of particle pnum - 1 sets n = pnum, else the
 * last lattice point sets n = sx sy sz (kernels/common.cl:535-559). */
void orc_populate_cuboid(real *pos, unsigned pnum, unsigned *n, const real *c1, const real *c2, unsigned sx,
                         unsigned sy, unsigned sz) {
    unsigned i, j, k;
    for (k = 0; k < sz; ++k)
        for (j = 0; j < sy; ++j)
            for (i = 0; i < sx; ++i) {
                const unsigned p = i + sx * j + sx * sy * k;
                vec v;
                if (p >= pnum) continue;
                if (p == pnum - 1) *n = pnum;
                else if (i == sx - 1 && j == sy - 1 && k == sz - 1) *n = sx * sy * sz;
                v.x = ((real)i / (real)(sx - 1)) * (c2[0] - c1[0]) + c1[0];
                v.y = ((real)j / (real)(sy - 1)) * (c2[1] - c1[1]) + c1[1];
                v.z = ((real)k / (real)(sz - 1)) * (c2[2] - c1[2]) + c1[2];
                st(pos, p, v);
            }
}

/* set_body_position: particles of body b (body_number == b + 1, b below
 * body_count) move by new - old and get velocity (new - old) / dt; the body's
 * stored position becomes new (kernels/rigidbody.cl:1-31).  Every particle
 * reads the old body position. */
void orc_set_body_position(real *pos, real *vel, const unsigned *body_number, unsigned n, unsigned body_count,
                           real dt, unsigned body, real *body_position, const real *newp) {
    const vec old = ld(body_position, body);
    const vec nw = {newp[0], newp[1], newp[2]};
    unsigned i;
    if (body < body_count)
        for (i = 0; i < n; ++i) {
            if (body_number[i] != body + 1) continue;
            st(pos, i, vadd(vsub(ld(pos, i), old), nw));
            st(vel, i, vdivs(vsub(nw, old), dt));
        }
    st(body_position, body, nw);
}

/* set_body_rotation: with B the matrix of the body's stored Euler angles and
 * P a particle's rotation field, T = B^T P; the position becomes T x and the
 * rotation field T P; the stored angles become the new ones
 * (kernels/rigidbody.cl:68-120). */
void orc_set_body_rotation(real *pos, real *rot, const unsigned *body_number, unsigned n, unsigned body_count,
                           unsigned body, real *body_euler, const real *euler_new) {
    real b[9], bt[9];
    unsigned i;
    int r, c;
    orc_euler_matrix(body_euler + 3 * (size_t)body, b);
    for (r = 0; r < 3; ++r)
        for (c = 0; c < 3; ++c) bt[3 * r + c] = b[3 * c + r];
    if (body < body_count)
        for (i = 0; i < n; ++i) {
            real t[9], t2[9];
            real *p = rot + 9 * (size_t)i;
            if (body_number[i] != body + 1) continue;
            mm3(bt, p, t);
            st(pos, i, mat_mv(t, ld(pos, i)));
            mm3(t, p, t2);
            memcpy(p, t2, sizeof t2);
        }
    body_euler[3 * (size_t)body] = euler_new[0];
    body_euler[3 * (size_t)body + 1] = euler_new[1];
    body_euler[3 * (size_t)body + 2] = euler_new[2];
}

unsigned orc_real_size(void) { return (unsigned)sizeof(real); }

#ifdef _OPENMP
#include <omp.h>
#endif
/* Threads the OpenMP passes use (1 without OpenMP). */
unsigned orc_threads(void) {
#ifdef _OPENMP
    return (unsigned)omp_get_max_threads();
#else
    return 1u;
#endif
}
