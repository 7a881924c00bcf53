// Solid passes: corotational linear elasticity and St. Venant-Kirchhoff.
//
// Reference kernels: compute_rotations_and_strains, compute_stresses,
// compute_forces_solids (kernels/solid_corotation.cl:28-403) and
// compute_deformations, compute_stresses, compute_forces_solids
// (kernels/solid_svk.cl:1-367), with the 3x3 helpers of
// kernels/common.cl:95-350.
//
// The polar decomposition (corotational) and the inverse of A_qq (SVK) run
// in FP64 when StepParams::polar_fp64 is set, in FP32 otherwise; both use the
// reference's fixed 8-rotation Jacobi.  The SVK Cauchy stress keeps all nine
// entries (the reference stores the 3x3 product into a 6-float array and
// reads sigma_zz past its end, kernels/solid_svk.cl:67,75).
#include "neighbours.h"

#pragma clang fp contract(off)

namespace sphk {
namespace {

constexpr unsigned kBlock = 256;
// List entries whose neighbour loads are issued together: the rotation and
// deformation sweeps load two positions per entry, the force pass ~27 floats.
constexpr unsigned kSweepGroup = 8;
constexpr unsigned kForceGroup = 4;
constexpr unsigned kForceGroupCoro = 3;   // + 9 rotation floats per entry: 3 keeps 3 waves per SIMD

template <bool KAHAN>
struct Acc {
    float s = 0.0f;
    __device__ __forceinline__ void add(float v) { s = s + v; }
    __device__ __forceinline__ float value() const { return s; }
};
template <>
struct Acc<true> {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float v) {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s; }
};

template <bool KAHAN>
struct Acc3 {
    Acc<KAHAN> x, y, z;
    __device__ __forceinline__ void add(V3 v) { x.add(v.x); y.add(v.y); z.add(v.z); }
    __device__ __forceinline__ V3 value() const { return V3{x.value(), y.value(), z.value()}; }
};

// apq += (w p) (x) q   (kernels/solids.cl:66-72)
template <bool KAHAN>
__device__ __forceinline__ void add_apq(Acc<KAHAN> *apq, V3 pv, V3 q, float w) {
    const V3 wp = w * pv;
    apq[0].add(wp.x * q.x); apq[1].add(wp.x * q.y); apq[2].add(wp.x * q.z);
    apq[3].add(wp.y * q.x); apq[4].add(wp.y * q.y); apq[5].add(wp.y * q.z);
    apq[6].add(wp.z * q.x); apq[7].add(wp.z * q.y); apq[8].add(wp.z * q.z);
}
// aqq += (w q) (x) q, upper triangle   (kernels/solids.cl:74-80)
template <bool KAHAN>
__device__ __forceinline__ void add_aqq(Acc<KAHAN> *aqq, V3 q, float w) {
    const V3 wq = w * q;
    aqq[S00].add(wq.x * q.x); aqq[S01].add(wq.x * q.y); aqq[S02].add(wq.x * q.z);
    aqq[S11].add(wq.y * q.y); aqq[S12].add(wq.y * q.z);
    aqq[S22].add(wq.z * q.z);
}
// out += a (x) b   (kernels/common.cl:252-256)
template <bool KAHAN>
__device__ __forceinline__ void add_outer(Acc<KAHAN> *out, V3 a, V3 b) {
    out[0].add(a.x * b.x); out[1].add(a.x * b.y); out[2].add(a.x * b.z);
    out[3].add(a.y * b.x); out[4].add(a.y * b.y); out[5].add(a.y * b.z);
    out[6].add(a.z * b.x); out[7].add(a.z * b.y); out[8].add(a.z * b.z);
}

// R = A (A^T A)^(-1/2)   (getR, kernels/common.cl:306-331); eigenvalue
// inverse square roots as 1/sqrt.
template <typename T>
__device__ __forceinline__ void polar_rotation(const float *apq, float *rot) {
    T a[9];
    for (int k = 0; k < 9; ++k) a[k] = static_cast<T>(apq[k]);
    T s[6] = {a[0] * a[0] + a[3] * a[3] + a[6] * a[6], a[1] * a[1] + a[4] * a[4] + a[7] * a[7],
              a[2] * a[2] + a[5] * a[5] + a[8] * a[8], a[0] * a[1] + a[3] * a[4] + a[6] * a[7],
              a[0] * a[2] + a[3] * a[5] + a[6] * a[8], a[1] * a[2] + a[4] * a[5] + a[7] * a[8]};
    T ev[3], pv[9], tmp[6], r[9];
    jacobi8(s, ev, pv);
    for (int k = 0; k < 3; ++k) ev[k] = T(1) / t_sqrt(ev[k]);
    p_d_pt(ev, pv, tmp);
    mul_full_sym(a, tmp, r);
    for (int k = 0; k < 9; ++k) rot[k] = static_cast<float>(r[k]);
}

// inverse of symmetric m by Jacobi (jacobiInvert, kernels/common.cl:296-305);
// in FP32 the reciprocals are taken in double as the 1.0 literal does there.
__device__ __forceinline__ void recip_ev(float *ev) {
    for (int k = 0; k < 3; ++k) ev[k] = static_cast<float>(1.0 / static_cast<double>(ev[k]));
}
__device__ __forceinline__ void recip_ev(double *ev) {
    for (int k = 0; k < 3; ++k) ev[k] = 1.0 / ev[k];
}
template <typename T>
__device__ __forceinline__ void deformation_from(const float *apq, const float *aqq, float *def) {
    T m[6], ev[3], pv[9], inv[6], a[9], out[9];
    for (int k = 0; k < 6; ++k) m[k] = static_cast<T>(aqq[k]);
    jacobi8(m, ev, pv);
    recip_ev(ev);
    p_d_pt(ev, pv, inv);
    for (int k = 0; k < 9; ++k) a[k] = static_cast<T>(apq[k]);
    mul_full_sym(a, inv, out);
    for (int k = 0; k < 9; ++k) def[k] = static_cast<float>(out[k]);
}

// Which lanes a launch of a solid neighbour pass handles.  With a kSolidList
// cache the pass runs as two kernels, one over the lanes whose list is
// complete and one scanning the candidates of the others (overflow, mirror
// planes: the density pass marks those lanes as overflowed), so that each
// kernel has only its own loop's register needs.  Without a cache one kernel
// scans every lane.
enum ListMode { kScanAll = 0, kListedOnly = 1, kScanOverflow = 2 };

// Entries of slot i's kSolidList list (> kListCap: scan instead).
template <int MODE>
__device__ __forceinline__ unsigned solid_list_count(const PairCache &c, unsigned i) {
    return MODE == kScanAll ? kListCap + 1u : c.count[i];
}
template <int MODE>
__device__ __forceinline__ bool skip_lane(unsigned cnt) {
    return (MODE == kListedOnly && cnt > kListCap) || (MODE == kScanOverflow && cnt <= kListCap);
}
__device__ __forceinline__ unsigned list_entry(const PairCache &c, unsigned k, unsigned i) {
    return c.list[static_cast<size_t>(k) * c.stride + i];
}

__device__ __forceinline__ bool in_body(const FieldPtrs &f, unsigned i) {
    return f.body_number != nullptr && f.body_number[i] > 0u;
}

template <class PS, bool KAHAN, bool FP64, int MODE>
__global__ void __launch_bounds__(kBlock) k_rotations_strains(PS ps, FieldPtrs f, PairCache c) {
    const unsigned i = ps.slot_xcd(blockDim.x);
    if (!ps.active(i) || in_body(f, i)) return;
    const unsigned cnt = solid_list_count<MODE>(c, i);
    if (skip_lane<MODE>(cnt)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float r2o = p.r2_out;   // read once: with a scene table p is in global memory
    const float *__restrict__ pos = f.position;
    const float *__restrict__ opos = f.originalpos;
    const V3 xi = load3(pos, i), x0i = load3(opos, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned ncombo = 1u << p.nplanes;
    const unsigned key = f.gridcell[i];

    Acc<KAHAN> apq[9];
    struct Pos2 {
        V3 x, x0;
    };
    auto load_pos2 = [&](unsigned e) {
        const unsigned j = e & kListSlot;
        return Pos2{load3(pos, j), load3(opos, j)};
    };
    if (MODE == kListedOnly) {
        for_each_listed<kSweepGroup>(c, i, cnt, load_pos2, [&](unsigned e, const Pos2 &d) {
            if (!(e & kListCurrent) || (e & kListSlot) == i) return;
            const V3 pv = d.x - xi;
            add_apq(apq, pv, d.x0 - x0i, w_poly6(length(pv), kc));
        });
    } else for_each_candidate(p.grid, p.p0, p.p0 + p.n, key, f.celloffset, f.gridcount, [&](unsigned j) {
        const V3 xj = load3(pos, j), x0j = load3(opos, j);
        if (j != i) {
            const V3 pv = xj - xi;
            const float s = sqlength(pv);   // r <= h as s < r2_out: no square root out of range
            if (!(s >= r2o)) add_apq(apq, pv, x0j - x0i, w_poly6(sqrtf(s), kc));
        }
        if (inear) {
            const unsigned ok = planes_near_j(p, xj, inear);
            if (ok)
                for (unsigned b = 1; b < ncombo; ++b) {
                    if (b & ~ok) continue;
                    const V3 pv = reflect_combo(p, b, xj) - xi;
                    const V3 q = reflect_combo(p, b, x0j) - x0i;
                    const float d = length(pv);
                    if (d < kc.h) add_apq(apq, pv, q, w_poly6(d, kc));
                }
        }
    });
    float a[9];
    for (int k = 0; k < 9; ++k) a[k] = apq[k].value();
    if (f.apq_temp)
        for (int k = 0; k < 9; ++k) f.apq_temp[9 * i + k] = a[k];
    if (f.intermediate)
        for (int k = 0; k < 9; ++k) f.intermediate[9 * i + k] = a[k];
    float rot[9];
    if (FP64) polar_rotation<double>(a, rot);
    else polar_rotation<float>(a, rot);
    for (int k = 0; k < 9; ++k) f.rotation[9 * i + k] = rot[k];
    const float rt[9] = {rot[0], rot[3], rot[6], rot[1], rot[4], rot[7], rot[2], rot[5], rot[8]};

    const float d0i = f.density0[i];
    const double rd0i = recip(d0i);   // the quotients (u * mass) / d0i below, bit for bit
    const float mass = p.mass;   // read once (see k_forces_solids)
    Acc<KAHAN> du[9];
    if (MODE == kListedOnly) {
        for_each_listed<kSweepGroup>(c, i, cnt, load_pos2, [&](unsigned e, const Pos2 &d) {
            if (!(e & kListReference) || (e & kListSlot) == i) return;
            const V3 q = d.x0 - x0i;
            const V3 u = mat_mul(rt, d.x - xi) - q;
            add_outer(du, div_recip(u * mass, rd0i), grad_spiky(-q, kc));
        });
    } else for_each_candidate(p.grid, p.p0, p.p0 + p.n, key, f.celloffset, f.gridcount, [&](unsigned j) {
        const V3 xj = load3(pos, j), x0j = load3(opos, j);
        if (j != i) {
            const V3 q = x0j - x0i;
            if (!(sqlength(q) >= r2o)) {
                const V3 u = mat_mul(rt, xj - xi) - q;
                add_outer(du, div_recip(u * mass, rd0i), grad_spiky(-q, kc));
            }
        }
        if (inear) {
            const unsigned ok = planes_near_j(p, xj, inear);
            if (ok)
                for (unsigned b = 1; b < ncombo; ++b) {
                    if (b & ~ok) continue;
                    const V3 pv = mat_mul(rt, reflect_combo(p, b, xj) - xi);
                    const V3 q = reflect_combo(p, b, x0j) - x0i;
                    const V3 u = pv - q;
                    if (length(pv) < kc.h) add_outer(du, div_recip(u * mass, rd0i), grad_spiky(-q, kc));
                }
        }
    });
    float g[9];
    for (int k = 0; k < 9; ++k) g[k] = du[k].value();
    float *st = f.strain + 6 * i;   // symmetric part (kernels/solid_corotation.cl:1-8)
    st[S00] = g[0];
    st[S11] = g[4];
    st[S22] = g[8];
    st[S01] = 0.5f * (g[1] + g[3]);
    st[S02] = 0.5f * (g[2] + g[6]);
    st[S12] = 0.5f * (g[5] + g[7]);
}

// sigma = C : eps, isotropic   (kernels/solid_corotation.cl:9-26,197-208)
template <class PS>
__global__ void __launch_bounds__(kBlock) k_stresses_corot(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_linear(kBlock);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const float E = p.youngs, nu = p.poisson;
    const float lam = E * nu / (1.0f + nu) / (1.0f - 2.0f * nu);
    const float mu = E / 2.0f / (1.0f + nu);
    const float d = lam + 2.0f * mu, sh = 2.0f * mu;
    const float c[6][6] = {{d, lam, lam, 0, 0, 0}, {lam, d, lam, 0, 0, 0}, {lam, lam, d, 0, 0, 0},
                           {0, 0, 0, sh, 0, 0},    {0, 0, 0, 0, sh, 0},    {0, 0, 0, 0, 0, sh}};
    const float *e = f.strain + 6 * i;
    float *s = f.stress + 6 * i;
    float ev[6];
    for (int k = 0; k < 6; ++k) ev[k] = e[k];
    for (int r = 0; r < 6; ++r)
        s[r] = c[r][0] * ev[0] + c[r][1] * ev[1] + c[r][2] * ev[2] + c[r][3] * ev[3] + c[r][4] * ev[4] +
               c[r][5] * ev[5];
}

template <class PS, bool KAHAN, bool FP64, int MODE>
__global__ void __launch_bounds__(kBlock) k_deformations_svk(PS ps, FieldPtrs f, PairCache c) {
    const unsigned i = ps.slot_xcd(blockDim.x);
    if (!ps.active(i)) return;
    const unsigned cnt = solid_list_count<MODE>(c, i);
    if (skip_lane<MODE>(cnt)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float r2o = p.r2_out;   // read once: with a scene table p is in global memory
    const float *__restrict__ pos = f.position;
    const float *__restrict__ opos = f.originalpos;
    const V3 xi = load3(pos, i), x0i = load3(opos, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned ncombo = 1u << p.nplanes;
    Acc<KAHAN> apq[9], aqq[6];
    if (MODE == kListedOnly) {
        struct Pos2 {
            V3 x, x0;
        };
        for_each_listed<kSweepGroup>(
            c, i, cnt,
            [&](unsigned e) {
                const unsigned j = e & kListSlot;
                return Pos2{load3(pos, j), load3(opos, j)};
            },
            [&](unsigned e, const Pos2 &d) {
                if (!(e & kListCurrent) || (e & kListSlot) == i) return;
                const V3 pv = d.x - xi;
                const V3 q = d.x0 - x0i;
                const float w = w_poly6(length(pv), kc);
                add_apq(apq, pv, q, w);
                add_aqq(aqq, q, w);
            });
    } else for_each_candidate(p.grid, p.p0, p.p0 + p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
        const V3 xj = load3(pos, j), x0j = load3(opos, j);
        if (j != i) {
            const V3 pv = xj - xi;
            const float s = sqlength(pv);
            if (!(s >= r2o)) {
                const V3 q = x0j - x0i;
                const float w = w_poly6(sqrtf(s), kc);
                add_apq(apq, pv, q, w);
                add_aqq(aqq, q, w);
            }
        }
        if (inear) {
            const unsigned ok = planes_near_j(p, xj, inear);
            if (ok)
                for (unsigned b = 1; b < ncombo; ++b) {
                    if (b & ~ok) continue;
                    const V3 pv = reflect_combo(p, b, xj) - xi;
                    const V3 q = reflect_combo(p, b, x0j) - x0i;
                    const float d = length(pv);
                    if (d < kc.h) {
                        const float w = w_poly6(d, kc);
                        add_apq(apq, pv, q, w);
                        add_aqq(aqq, q, w);
                    }
                }
        }
    });
    float a[9], m[6], def[9];
    for (int k = 0; k < 9; ++k) a[k] = apq[k].value();
    for (int k = 0; k < 6; ++k) m[k] = aqq[k].value();
    if (FP64) deformation_from<double>(a, m, def);
    else deformation_from<float>(a, m, def);
    for (int k = 0; k < 9; ++k) f.deformation[9 * i + k] = def[k];
}

// Green strain -> SVK 2nd Piola-Kirchhoff -> Cauchy, with the double
// literals of kernels/solid_svk.cl:1-43 kept as double arithmetic.
template <class PS>
__global__ void __launch_bounds__(kBlock) k_stresses_svk(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_linear(kBlock);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    float F[9], Ft[9], ftf[9];
    for (int k = 0; k < 9; ++k) F[k] = f.deformation[9 * i + k];
    Ft[0] = F[0]; Ft[4] = F[4]; Ft[8] = F[8];
    Ft[1] = F[3]; Ft[3] = F[1]; Ft[2] = F[6]; Ft[6] = F[2]; Ft[5] = F[7]; Ft[7] = F[5];
    mul_full(Ft, F, ftf);
    float e[6];
    e[S00] = static_cast<float>(0.5 * ftf[0] - 0.5);
    e[S11] = static_cast<float>(0.5 * ftf[4] - 0.5);
    e[S22] = static_cast<float>(0.5 * ftf[8] - 0.5);
    e[S01] = static_cast<float>(0.5 * ftf[1]);
    e[S02] = static_cast<float>(0.5 * ftf[2]);
    e[S12] = static_cast<float>(0.5 * ftf[5]);
    const float E = p.youngs, nu = p.poisson;
    const float lam = static_cast<float>(static_cast<double>(E * nu) / (1.0 + nu) / (1.0 - 2.0 * nu));
    const float mu = static_cast<float>(E / 2.0 / (1.0 + nu));
    const float first = lam * (e[S00] + e[S11] + e[S22]);
    float svk[6];
    svk[S00] = static_cast<float>(first + 2.0 * mu * e[S00]);
    svk[S11] = static_cast<float>(first + 2.0 * mu * e[S11]);
    svk[S22] = static_cast<float>(first + 2.0 * mu * e[S22]);
    svk[S01] = static_cast<float>(2.0 * mu * e[S01]);
    svk[S02] = static_cast<float>(2.0 * mu * e[S02]);
    svk[S12] = static_cast<float>(2.0 * mu * e[S12]);
    float tmp[9], cauchy[9];
    mul_full_sym(F, svk, tmp);
    mul_full(tmp, Ft, cauchy);
    const float sc = static_cast<float>(1.0 / det3(F));
    float *s = f.stress + 6 * i;
    s[S00] = sc * cauchy[0];
    s[S11] = sc * cauchy[4];
    // svk_sigma_zz_zero: the value the reference stores (its sigma_zz is read
    // past the end of a float[6], kernels/solid_svk.cl:48,75, and comes out 0)
    s[S22] = p.svk_szz_zero ? 0.0f : sc * cauchy[8];
    s[S01] = sc * cauchy[1];
    s[S02] = sc * cauchy[2];
    s[S12] = sc * cauchy[5];
}

template <class PS, bool CORO, bool KAHAN, int MODE>
__global__ void __launch_bounds__(kBlock) k_forces_solids(PS ps, FieldPtrs f, PairCache c) {
    const unsigned i = ps.slot_xcd(blockDim.x);
    if (!ps.active(i)) return;
    const unsigned cnt = solid_list_count<MODE>(c, i);
    if (skip_lane<MODE>(cnt)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float r2o = p.r2_out;   // read once: with a scene table p is in global memory
    const float *__restrict__ pos = f.position;
    const float *__restrict__ opos = f.originalpos;
    const float *__restrict__ vel = f.velocity;
    const float *__restrict__ stress = f.stress;
    const float *__restrict__ rot = f.rotation;
    const V3 xi = load3(pos, i), vi = load3(vel, i), x0i = load3(opos, i);
    float si[6], ri[9];
    for (int k = 0; k < 6; ++k) si[k] = stress[6 * i + k];
    if (CORO)
        for (int k = 0; k < 9; ++k) ri[k] = rot[9 * i + k];
    const float d0i = f.density0[i];
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned ncombo = 1u << p.nplanes;
    // Scene scalars read once: with a scene table `p` lives in global memory
    // and the pair loops would otherwise reload it per pair.
    const float mass = p.mass;
    Acc3<KAHAN> fe, fv, fec, fvc;
    // One in-range pair, direct term (kernels/solid_corotation.cl:262-300).
    auto pair = [&](unsigned j, V3 xj, V3 x0j, V3 vj, float vivj, const float *sj, const float *rj, float mrho) {
        const V3 X = x0i - x0j;
        const float rcur = length(xj - xi);
        V3 g, gm;   // grad_spiky(X), grad_spiky(-X)
        grad_spiky_pm(X, kc, &g, &gm);
        const V3 fji = -vivj * sym_mul(si, g);
        const V3 fij = (f.body_number && load1(f.body_number, j) > 0u) ? -fji : -vivj * sym_mul(sj, gm);
        const V3 e = CORO ? (-mat_mul(ri, fji)) + mat_mul(rj, fij) : (-fji) + fij;
        const V3 v = (mrho * (vj - vi)) * lap_visc(rcur, kc);
        fe.add(e);
        fv.add(v);
        if (MODE != kListedOnly) {   // listed lanes have no mirror images: cutoff sums = full sums
            fec.add(e);
            fvc.add(v);
        }
    };
    if (MODE == kListedOnly) {
        struct Nbr {
            V3 x, x0, v;
            float d0, rho, s[6], r[CORO ? 9 : 1];
        };
        for_each_listed<(CORO ? kForceGroupCoro : kForceGroup)>(
            c, i, cnt,
            [&](unsigned e) {
                const unsigned j = e & kListSlot;
                Nbr n;
                n.x = load3(pos, j);
                n.x0 = load3(opos, j);
                n.v = load3(vel, j);
                n.d0 = load1(f.density0, j);
                n.rho = load1(f.density, j);
                const float *sj = at_offset(stress, j * 24u);
                for (int m = 0; m < 6; ++m) n.s[m] = sj[m];
                if (CORO) {
                    const float *rj = at_offset(rot, j * 36u);
                    for (int m = 0; m < 9; ++m) n.r[m] = rj[m];
                }
                return n;
            },
            [&](unsigned e, const Nbr &n) {
                const unsigned j = e & kListSlot;
                if (j == i) return;
                const float vivj = mass * mass / d0i / n.d0;
                pair(j, n.x, n.x0, n.v, vivj, n.s, CORO ? n.r : nullptr, mass / n.rho);
            });
    } else for_each_candidate(p.grid, p.p0, p.p0 + p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
        const V3 x0j = load3(opos, j);
        const V3 xj = load3(pos, j);
        const V3 X = x0i - x0j;
        const bool near_pair = !(sqlength(X) >= r2o) || !(sqlength(xj - xi) >= r2o);
        if (!near_pair && !inear) return;
        const V3 vj = load3(vel, j);
        const float vivj = mass * mass / d0i / f.density0[j];
        float sj[6], rj[9];
        for (int k = 0; k < 6; ++k) sj[k] = stress[6 * j + k];
        if (CORO)
            for (int k = 0; k < 9; ++k) rj[k] = rot[9 * j + k];
        const float mrho = mass / f.density[j];
        if (j != i && near_pair) pair(j, xj, x0j, vj, vivj, sj, rj, mrho);
        if (inear) {
            const unsigned ok = planes_near_j(p, xj, inear);
            if (ok)
                for (unsigned b = 1; b < ncombo; ++b) {
                    if (b & ~ok) continue;
                    const V3 rp = reflect_combo(p, b, xj);
                    const V3 rp0 = reflect_combo(p, b, x0j);
                    const V3 rip0 = reflect_combo(p, b, x0i);
                    const V3 Xm = x0i - rp0;
                    const V3 refx = rip0 - x0j;
                    const V3 fji = -vivj * sym_mul(si, grad_spiky(Xm, kc));
                    V3 fij = -vivj * sym_mul(sj, grad_spiky(-refx, kc));
                    if (CORO) fij = mat_mul(rj, fij);
                    unsigned lowest = 0;
                    for (unsigned s = p.nplanes; s > 0; --s)
                        if ((b >> (s - 1)) & 1u) {
                            fij = reflect_dir(fij, plane_normal(p, s - 1));
                            lowest = s - 1;
                        }
                    const V3 e = CORO ? (-mat_mul(ri, fji)) + fij : (-fji) + fij;
                    fe.add(e);
                    const V3 refvel = reflect_dir(vj, plane_normal(p, lowest));
                    const V3 v = (mrho * (refvel - vi)) * lap_visc(length(rp - xi), kc);
                    fv.add(v);
                    if (!(b & 1u)) {
                        fec.add(e);
                        fvc.add(v);
                    }
                }
        }
    });
    const V3 e = fe.value() * 0.5f;
    const V3 v = fv.value() * p.viscosity;
    store3(f.force, i, e + v);
    if (f.force_elastic) store3(f.force_elastic, i, e);
    if (f.force_viscous) store3(f.force_viscous, i, v);
    if (f.cutoff_force) {
        const V3 ec = MODE == kListedOnly ? fe.value() : fec.value();
        const V3 vc = MODE == kListedOnly ? fv.value() : fvc.value();
        store3(f.cutoff_force, i, ec * 0.5f + vc * p.viscosity);
    }
}

// Launches kernel template K<MODE> as the list-only and overflow-scan pair
// when `c` holds a solid list, else as one scan over every lane, with `nb`
// threads per block.
template <template <int> class K, class PS>
hipError_t launch_modes(const PS &ps, const FieldPtrs &f, const PairCache &c, hipStream_t s, unsigned nb) {
    if (!ps.has_work()) return hipSuccess;
    const dim3 g = ps.launch_blocks(nb, true), b(nb);
    if (c.valid == kSolidList) {
        hipLaunchKernelGGL(K<kListedOnly>::template kernel<PS>, g, b, 0, s, ps, f, c);
        hipLaunchKernelGGL(K<kScanOverflow>::template kernel<PS>, g, b, 0, s, ps, f, c);
    } else {
        hipLaunchKernelGGL(K<kScanAll>::template kernel<PS>, g, b, 0, s, ps, f, c);
    }
    return hipGetLastError();
}

template <bool KAHAN, bool FP64>
struct RotationsFor {
    template <int MODE>
    struct M {
        template <class PS>
        static constexpr auto kernel = k_rotations_strains<PS, KAHAN, FP64, MODE>;
    };
};
template <bool KAHAN, bool FP64>
struct DeformationsFor {
    template <int MODE>
    struct M {
        template <class PS>
        static constexpr auto kernel = k_deformations_svk<PS, KAHAN, FP64, MODE>;
    };
};
template <bool CORO, bool KAHAN>
struct ForcesFor {
    template <int MODE>
    struct M {
        template <class PS>
        static constexpr auto kernel = k_forces_solids<PS, CORO, KAHAN, MODE>;
    };
};

// The rotation/strain and deformation passes keep their per-lane 3x3 work
// arrays in LDS, sized by the compiler for kBlock threads per block whatever
// the launch size, so they always run kBlock-thread blocks: smaller blocks
// would leave that LDS idle and cap the waves per CU (64-thread blocks: 2 per
// SIMD).
template <class PS>
hipError_t rotations_and_strains(const PS &ps, const FieldPtrs &f, const PairCache &c, int compensated,
                                 int polar_fp64, hipStream_t s) {
    const bool k = (compensated & kKahanForces) != 0;
    if (k && polar_fp64) return launch_modes<RotationsFor<true, true>::template M>(ps, f, c, s, kBlock);
    if (k) return launch_modes<RotationsFor<true, false>::template M>(ps, f, c, s, kBlock);
    if (polar_fp64) return launch_modes<RotationsFor<false, true>::template M>(ps, f, c, s, kBlock);
    return launch_modes<RotationsFor<false, false>::template M>(ps, f, c, s, kBlock);
}

template <class PS>
hipError_t stresses_corotational(const PS &ps, const FieldPtrs &f, hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    hipLaunchKernelGGL((k_stresses_corot<PS>), ps.launch_blocks(kBlock, false), dim3(kBlock), 0, s, ps, f);
    return hipGetLastError();
}

template <class PS>
hipError_t deformations_svk(const PS &ps, const FieldPtrs &f, const PairCache &c, int compensated, int polar_fp64,
                            hipStream_t s) {
    const bool k = (compensated & kKahanForces) != 0;
    if (k && polar_fp64) return launch_modes<DeformationsFor<true, true>::template M>(ps, f, c, s, kBlock);
    if (k) return launch_modes<DeformationsFor<true, false>::template M>(ps, f, c, s, kBlock);
    if (polar_fp64) return launch_modes<DeformationsFor<false, true>::template M>(ps, f, c, s, kBlock);
    return launch_modes<DeformationsFor<false, false>::template M>(ps, f, c, s, kBlock);
}

template <class PS>
hipError_t stresses_svk(const PS &ps, const FieldPtrs &f, hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    hipLaunchKernelGGL((k_stresses_svk<PS>), ps.launch_blocks(kBlock, false), dim3(kBlock), 0, s, ps, f);
    return hipGetLastError();
}

template <class PS>
hipError_t forces_solids(const PS &ps, const FieldPtrs &f, const PairCache &c, bool corotational, int compensated,
                         hipStream_t s) {
    const bool k = (compensated & kKahanForces) != 0;
    const unsigned nb = neighbour_block();
    if (corotational && k) return launch_modes<ForcesFor<true, true>::template M>(ps, f, c, s, nb);
    if (corotational) return launch_modes<ForcesFor<true, false>::template M>(ps, f, c, s, nb);
    if (k) return launch_modes<ForcesFor<false, true>::template M>(ps, f, c, s, nb);
    return launch_modes<ForcesFor<false, false>::template M>(ps, f, c, s, nb);
}

}  // namespace

hipError_t launch_rotations_and_strains(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                        hipStream_t s) {
    return rotations_and_strains(SingleScene{p}, f, c, p.compensated, p.polar_fp64, s);
}

hipError_t launch_stresses_corotational(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    return stresses_corotational(SingleScene{p}, f, s);
}

hipError_t launch_deformations_svk(const StepParams &p, const FieldPtrs &f, const PairCache &c, hipStream_t s) {
    return deformations_svk(SingleScene{p}, f, c, p.compensated, p.polar_fp64, s);
}

hipError_t launch_stresses_svk(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    return stresses_svk(SingleScene{p}, f, s);
}

hipError_t launch_forces_solids(const StepParams &p, const FieldPtrs &f, const PairCache &c, bool corotational,
                                hipStream_t s) {
    return forces_solids(SingleScene{p}, f, c, corotational, p.compensated, s);
}

hipError_t launch_batch_rotations_and_strains(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                              const PairCache &c, int compensated, int polar_fp64, hipStream_t s) {
    return rotations_and_strains(SceneTable{b, sel}, f, c, compensated, polar_fp64, s);
}

hipError_t launch_batch_stresses_corotational(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                              hipStream_t s) {
    return stresses_corotational(SceneTable{b, sel}, f, s);
}

hipError_t launch_batch_deformations_svk(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                         const PairCache &c, int compensated, int polar_fp64, hipStream_t s) {
    return deformations_svk(SceneTable{b, sel}, f, c, compensated, polar_fp64, s);
}

hipError_t launch_batch_stresses_svk(const BatchTable &b, SceneFilter sel, const FieldPtrs &f, hipStream_t s) {
    return stresses_svk(SceneTable{b, sel}, f, s);
}

hipError_t launch_batch_forces_solids(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                      const PairCache &c, bool corotational, int compensated, hipStream_t s) {
    return forces_solids(SceneTable{b, sel}, f, c, corotational, compensated, s);
}

}  // namespace sphk
