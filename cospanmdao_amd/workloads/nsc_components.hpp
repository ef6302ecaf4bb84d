#pragma once

// =============================================================================
// cospanmdao_amd/workloads/nsc_components.hpp — the compressible Navier-Stokes residual components
// =============================================================================
//
// The reference's examples/ns-residual: one evaluation of R = -div F_conv + div F_visc per cell of a uniform 2D grid
// (conserved U = (rho, rho u, rho v, E) as [4, ny, nx] f32, ideal gas gamma 1.4, mu 0.01, Pr 0.72, dx = dy = 1), x periodic
// and y periodic or bounded by no-slip adiabatic walls (ghost rows mirror the interior with both momenta negated), MUSCL
// on primitives with a minmod or van Leer limiter, HLLC with Davis speeds, the viscous flux from the face's two cells
// (normal derivative) and the mean of their central-difference gradients (tangential derivative), heat flux k grad T.
// Diagnostics (the reference's README.md:5-14): the minimum local timestep, the maximum wave speed and the sum of R^2 per
// variable, each computed atomically (fetch_min / fetch_max / fetch_add on atomic_f32 outputs that an init stage resets)
// and deterministically (per-tile partials reduced in a fixed order), and an error stage writing atomic - deterministic.
// The scheme and its conventions are the reference's (physics.hpp:14-33): the face fluxes must be the same functions of
// the same cells for the residual to be the same.
//
// NscTile<Lim, Walls, TY> (the MI355X design of the tile pipeline, one stage): a work-group of 256 owns a tile of TY rows
// x 64 columns. It loads the tile with a 2-cell halo once (4 coalesced loads per cell, boundary rows mirrored or wrapped
// on the way) and converts it to primitives in LDS; computes EVERY face of its cells (the x-faces of the
// tile's east edge and the y-faces of its north edge too, which the neighbouring tile computes again: 1/64 and 1/TY more
// face work instead of any exchange between work-groups), x-faces first into an LDS flux buffer, their divergence into
// registers, then the y-faces into the same buffer; and finishes each cell: R to HBM, the cell's timestep and wave speed,
// the squares of R. The flux through a face, convective minus viscous, is one value per variable, so a cell's residual is
// two differences per direction. Tiles go to XCDs in contiguous runs (the halo rows of vertically adjacent tiles meet in one
// L2). Divisions are reciprocal multiplies and square roots the hardware's (v_rcp_f32, v_sqrt_f32; 1 ulp): the checks are
// the reference's (R within 1e-4 relative L2 of the FP64 residual) and every face is computed by the same instructions
// wherever it is computed, so a uniform periodic flow still gives R exactly 0.

#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>

#include "naga/cospan/cospan.hpp"

namespace nsc {

namespace cs = naga::cospan;
using cs::DType;

inline constexpr int kWg   = 256;
inline constexpr int kTX   = 64;  ///< tile columns
inline constexpr int kHalo = 2;

/// Limiters (physics.hpp:52-69): the slope of a cell from its backward difference a and forward difference b.
inline constexpr int kMinMod  = 0;
inline constexpr int kVanLeer = 1;

CS_COMPONENTS_BEGIN

namespace detail {

inline constexpr float kGamma  = 1.4f;
inline constexpr float kGm1    = kGamma - 1.0f;
inline constexpr float kInvGm1 = 1.0f / kGm1;
inline constexpr float kMu     = 0.01f;
inline constexpr float kPr     = 0.72f;
inline constexpr float kCfl    = 0.5f;
inline constexpr float kKappa  = kMu * kGamma / (kGm1 * kPr);                        ///< heat conductivity
inline constexpr float kNuMax  = (kGamma / kPr > 4.0f / 3.0f ? kGamma / kPr : 4.0f / 3.0f);  ///< max(4/3, gamma/Pr)
inline constexpr float kTwoThirds = 2.0f / 3.0f;

__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float root(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int Lim>
__device__ inline float slope(float a, float b) {
    const float ab = a * b;
    float       s;
    if constexpr (Lim == kMinMod)
        s = a > 0.0f ? fminf(a, b) : fmaxf(a, b);
    else
        s = 2.0f * ab * rcp(a + b);
    return ab <= 0.0f ? 0.0f : s;
}

/// HLLC through a face with normal x (Dir 0) or y (Dir 1) from the reconstructed primitives of the two sides.
template <int Dir>
__device__ inline void hllc(const float (&wl)[4], const float (&wr)[4], float (&F)[4]) {
    constexpr int N = 1 + Dir, T = 2 - Dir;  // normal and tangential velocity slots
    const float   rl = wl[0], rr = wr[0], pl = wl[3], pr = wr[3];
    const float   ul = wl[N], vl = wl[T], ur = wr[N], vr = wr[T];
    const float   irl = rcp(rl), irr = rcp(rr);
    const float   cl = root(kGamma * pl * irl), cr = root(kGamma * pr * irr);
    const float   El = pl * kInvGm1 + 0.5f * rl * (ul * ul + vl * vl);
    const float   Er = pr * kInvGm1 + 0.5f * rr * (ur * ur + vr * vr);
    const float   sl = fminf(ul - cl, ur - cr), sr = fmaxf(ul + cl, ur + cr);  // Davis
    const float   ml = rl * (sl - ul), mr = rr * (sr - ur);                   // mass flux into each star region
    const float   sm = (pr - pl + ul * ml - ur * mr) * rcp(ml - mr);
    // the side whose flux is taken: left when supersonic to the right, or subsonic with the contact moving right
    const bool  left = sl >= 0.0f || (sr > 0.0f && sm >= 0.0f);
    const bool  star = sl < 0.0f && sr > 0.0f;
    const float r = left ? rl : rr, un = left ? ul : ur, ut = left ? vl : vr, p = left ? pl : pr, E = left ? El : Er;
    const float ir = left ? irl : irr, m = left ? ml : mr, s = left ? sl : sr;
    const float f0 = r * un, f1 = r * un * un + p, f2 = r * un * ut, f3 = (E + p) * un;
    // the star state's correction F* = F + s (U* - U), U* = (k, k sm, k ut, Es)
    const float k  = m * rcp(s - sm);
    const float Es = k * (E * ir + (sm - un) * (sm + p * rcp(m)));
    const float g0 = f0 + s * (k - r), g1 = f1 + s * (k * sm - r * un), g2 = f2 + s * (k * ut - r * ut), g3 = f3 + s * (Es - E);
    F[0] = star ? g0 : f0;
    F[N] = star ? g1 : f1;
    F[T] = star ? g2 : f2;
    F[3] = star ? g3 : f3;
}

/// The viscous flux through a face with normal Dir: dn* the differences between the two cells along the normal (u, v, T),
/// tu / tv the tangential derivatives of u and v (mean of the cells' central differences), uf / vf the face velocity.
template <int Dir>
__device__ inline void viscous(float dnu, float dnv, float dnT, float tu, float tv, float uf, float vf, float (&F)[4]) {
    const float ux = Dir == 0 ? dnu : tu, uy = Dir == 0 ? tu : dnu;
    const float vx = Dir == 0 ? dnv : tv, vy = Dir == 0 ? tv : dnv;
    const float div = ux + vy;
    const float txx = kMu * (2.0f * ux - kTwoThirds * div), tyy = kMu * (2.0f * vy - kTwoThirds * div), txy = kMu * (uy + vx);
    const float a = Dir == 0 ? txx : txy, b = Dir == 0 ? txy : tyy;
    F[0]          = 0.0f;
    F[1]          = a;
    F[2]          = b;
    F[3]          = uf * a + vf * b + kKappa * dnT;
}

}  // namespace detail

/// The LDS of a tile: 4 fields (rho u v p) of (TY + 4) x 68 cells, then 4 flux fields (one direction at a time). The
/// temperatures of a face's two cells are recomputed (two reciprocals) rather than stored: at 16 rows that keeps a tile
/// under 40 KB, four work-groups per CU instead of three (the tile is VALU-bound and needs the waves to hide latency).
using TileMem = cs::LocalMem<cs::LocalArray<DType::f32, 1>>;

// =============================================================================
// NscTile: the residual, per-tile partials and the atomic diagnostics of a tile
// =============================================================================
template <int Lim, bool Walls, int TY>
class NscTile : public cs::Component<NscTile<Lim, Walls, TY>> {
public:
    static constexpr int max_work_group_size = kWg;
    static constexpr int PW = kTX + 2 * kHalo, PH = TY + 2 * kHalo, PC = PW * PH;
    static constexpr int NFX = TY * (kTX + 1), NFY = (TY + 1) * kTX;
    static constexpr int FS          = NFX > NFY ? NFX : NFY;  ///< one flux field
    static constexpr int kLocalElems = 4 * PC + 4 * FS;
    static constexpr int kCells      = TY * kTX / kWg;  ///< cells per work-item
    static_assert(TY * kTX % kWg == 0, "a tile is a whole number of cells per work-item");
    static_assert(kLocalElems * 4 <= 40 * 1024, "a tile's LDS lets four work-groups share a CU");

    struct Inputs {
        cs::Input<DType::f32, 3> U;  ///< [4, ny, nx] conserved
    };
    struct Outputs {
        cs::Output<DType::f32, 3>        R;     ///< [4, ny, nx]
        cs::Output<DType::f32, 2>        Part;  ///< [6, tiles]: min dt, max wave speed, sum R^2 of the 4 variables
        cs::Output<DType::atomic_f32, 1> Dt, Ws, Sq;
    };

    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out, TileMem lm) {
        using namespace detail;
        auto [lt]          = lm.get(it);
        float* const P     = &lt[std::size_t{0}];  // P + f * PC: rho, u, v, p
        float* const Fl    = P + 4 * PC;            // Fl + q * FS: the face fluxes of variable q
        const int    tid   = static_cast<int>(it.get_local_id(0));
        const auto   shape = in.U.shape();
        const int    ny = static_cast<int>(shape[1]), nx = static_cast<int>(shape[2]);
        const int    ntx = nx / kTX, tiles = ntx * (ny / TY);
        int          g   = static_cast<int>(it.get_group(0));
        if (tiles % 8 == 0) g = (g % 8) * (tiles / 8) + g / 8;  // XCD k runs tiles [k tiles / 8, (k + 1) tiles / 8)
        const int         y0 = (g / ntx) * TY, x0 = (g % ntx) * kTX;
        const std::size_t plane = static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny);
        const float*      U     = cs::raw::data(in.U);

        // ---- the tile and its halo: conserved -> primitives ----
        for (int e = tid; e < PC; e += kWg) {
            const int b = e / PW, a = e - b * PW;
            int       gx = x0 - kHalo + a;
            gx += gx < 0 ? nx : 0;
            gx -= gx >= nx ? nx : 0;
            int   gy   = y0 - kHalo + b;
            float flip = 1.0f;
            if (gy < 0 || gy >= ny) {
                if constexpr (Walls) {
                    gy   = gy < 0 ? -1 - gy : 2 * ny - 1 - gy;
                    flip = -1.0f;
                } else {
                    gy += gy < 0 ? ny : -ny;
                }
            }
            const std::size_t c  = static_cast<std::size_t>(gy) * static_cast<std::size_t>(nx) + static_cast<std::size_t>(gx);
            const float       r  = U[c];
            const float       ir = rcp(r);
            const float       u = flip * U[plane + c] * ir, v = flip * U[2 * plane + c] * ir;
            const float       p = kGm1 * (U[3 * plane + c] - 0.5f * r * (u * u + v * v));
            P[e]                = r;
            P[PC + e]           = u;
            P[2 * PC + e]       = v;
            P[3 * PC + e]       = p;
        }
        __syncthreads();
        const float* const Pu = P + PC;
        const float* const Pv = P + 2 * PC;
        const float* const Pp = P + 3 * PC;

        // ---- x-faces: face c of tile row r lies between cells (r, c - 1) and (r, c) ----
        for (int f = tid; f < NFX; f += kWg) {
            const int r = f / (kTX + 1), c = f - r * (kTX + 1);
            const int e = (r + kHalo) * PW + c + kHalo - 1;  // the left cell; the right one is e + 1
            float     wl[4], wr[4], Fc[4], Fv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float* F  = P + q * PC;
                const float  a0 = F[e - 1], a1 = F[e], a2 = F[e + 1], a3 = F[e + 2];
                wl[q]           = a1 + 0.5f * slope<Lim>(a1 - a0, a2 - a1);
                wr[q]           = a2 - 0.5f * slope<Lim>(a2 - a1, a3 - a2);
            }
            hllc<0>(wl, wr, Fc);
            const float tu = 0.5f * (0.5f * (Pu[e + PW] - Pu[e - PW]) + 0.5f * (Pu[e + 1 + PW] - Pu[e + 1 - PW]));
            const float tv = 0.5f * (0.5f * (Pv[e + PW] - Pv[e - PW]) + 0.5f * (Pv[e + 1 + PW] - Pv[e + 1 - PW]));
            viscous<0>(Pu[e + 1] - Pu[e], Pv[e + 1] - Pv[e], Pp[e + 1] * rcp(P[e + 1]) - Pp[e] * rcp(P[e]), tu, tv, 0.5f * (Pu[e] + Pu[e + 1]), 0.5f * (Pv[e] + Pv[e + 1]), Fv);
#pragma unroll
            for (int q = 0; q < 4; ++q) Fl[q * FS + f] = Fc[q] - Fv[q];
        }
        __syncthreads();
        float Rc[kCells][4];
#pragma unroll
        for (int k = 0; k < kCells; ++k) {
            const int cell = tid + k * kWg, r = cell / kTX, c = cell - r * kTX, f = r * (kTX + 1) + c;
#pragma unroll
            for (int q = 0; q < 4; ++q) Rc[k][q] = Fl[q * FS + f] - Fl[q * FS + f + 1];
        }
        __syncthreads();

        // ---- y-faces: face r of tile column c lies between cells (r - 1, c) and (r, c) ----
        for (int f = tid; f < NFY; f += kWg) {
            const int r = f / kTX, c = f - r * kTX;
            const int e = (r + kHalo - 1) * PW + c + kHalo;  // the cell below; the one above is e + PW
            float     wl[4], wr[4], Fc[4], Fv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float* F  = P + q * PC;
                const float  a0 = F[e - PW], a1 = F[e], a2 = F[e + PW], a3 = F[e + 2 * PW];
                wl[q]           = a1 + 0.5f * slope<Lim>(a1 - a0, a2 - a1);
                wr[q]           = a2 - 0.5f * slope<Lim>(a2 - a1, a3 - a2);
            }
            hllc<1>(wl, wr, Fc);
            const float tu = 0.5f * (0.5f * (Pu[e + 1] - Pu[e - 1]) + 0.5f * (Pu[e + PW + 1] - Pu[e + PW - 1]));
            const float tv = 0.5f * (0.5f * (Pv[e + 1] - Pv[e - 1]) + 0.5f * (Pv[e + PW + 1] - Pv[e + PW - 1]));
            viscous<1>(Pu[e + PW] - Pu[e], Pv[e + PW] - Pv[e], Pp[e + PW] * rcp(P[e + PW]) - Pp[e] * rcp(P[e]), tu, tv, 0.5f * (Pu[e] + Pu[e + PW]),
                       0.5f * (Pv[e] + Pv[e + PW]), Fv);
#pragma unroll
            for (int q = 0; q < 4; ++q) Fl[q * FS + f] = Fc[q] - Fv[q];
        }
        __syncthreads();

        // ---- the cells: R, the local timestep, the wave speed, the squares ----
        // R goes out with nontemporal stores: no stage of the run reads it back.
        float* const Rout = cs::raw::data(out.R);
        float        dt = std::numeric_limits<float>::infinity(), ws = -std::numeric_limits<float>::infinity();
        float        sq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < kCells; ++k) {
            const int         cell = tid + k * kWg, r = cell / kTX, c = cell - r * kTX;
            const std::size_t o    = static_cast<std::size_t>(y0 + r) * static_cast<std::size_t>(nx) + static_cast<std::size_t>(x0 + c);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float R = Rc[k][q] + (Fl[q * FS + r * kTX + c] - Fl[q * FS + (r + 1) * kTX + c]);
                __builtin_nontemporal_store(R, Rout + (q * plane + o));
                sq[q] += R * R;
            }
            const int   e  = (r + kHalo) * PW + c + kHalo;
            const float ir = rcp(P[e]);
            const float cs_ = root(kGamma * Pp[e] * ir);
            const float ax = fabsf(Pu[e]) + cs_, ay = fabsf(Pv[e]) + cs_;
            dt             = fminf(dt, kCfl * rcp(ax + ay + 4.0f * (kMu * ir * kNuMax)));
            ws             = fmaxf(ws, fmaxf(ax, ay));
        }
        const auto grp = it.get_group();
        dt             = cs::reduce_over_group(grp, dt, cs::minimum<float>());
        ws             = cs::reduce_over_group(grp, ws, cs::maximum<float>());
#pragma unroll
        for (int q = 0; q < 4; ++q) sq[q] = cs::reduce_over_group(grp, sq[q], cs::plus<float>());
        if (tid == 0) {
            float* const      part = cs::raw::data(out.Part);
            const std::size_t t    = static_cast<std::size_t>(g), n = static_cast<std::size_t>(tiles);
            part[t]                = dt;
            part[n + t]            = ws;
#pragma unroll
            for (int q = 0; q < 4; ++q) part[(2 + q) * n + t] = sq[q];
        }
        // the atomics after the init stage (in a fused group with waits it is group 0 of segment 0; elsewhere it ran in an
        // earlier launch and the wait is a no-op). Device-scope atomics on one address serialize across the XCDs: a min
        // (max) is only sent when the tile's value is below (above) the accumulator's current value, which only decreases
        // (increases) after the reset, so the result is the same; the four sums go out from four work-items at once.
        it.wait_group(0, cs::id<1>{0});
        if (tid == 0) {
            const auto a = out.Dt[cs::id<1>{0}];
            if (dt < a.load()) a.fetch_min(dt);
        } else if (tid == 1) {
            const auto a = out.Ws[cs::id<1>{0}];
            if (ws > a.load()) a.fetch_max(ws);
        } else if (tid < 6) {
            const int q = tid - 2;
            out.Sq[cs::id<1>{static_cast<std::size_t>(q)}].fetch_add(q == 0 ? sq[0] : q == 1 ? sq[1] : q == 2 ? sq[2] : sq[3]);
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        this->declare_input(inputs.U, "U", cs::unspecified_shape<3>());
        this->declare_output(outputs.R, "R", cs::unspecified_shape<3>());
        this->declare_output(outputs.Part, "Part", cs::unspecified_shape<2>());
        this->declare_output(outputs.Dt, "Dt", 1);
        this->declare_output(outputs.Ws, "Ws", 1);
        this->declare_output(outputs.Sq, "Sq", 4);
    }
    void prepare_forward() override {
        const cs::range<3> s = this->shape_of(inputs.U);
        if (s[0] != 4 || s[2] % kTX != 0 || s[1] % TY != 0 || s[2] == 0 || s[1] == 0)
            throw std::invalid_argument("nsc: U must be [4, ny, nx] with nx a multiple of " + std::to_string(kTX) + " and ny a multiple of " +
                                        std::to_string(TY));
        this->set_shape(outputs.R, s[0], s[1], s[2]);
        this->set_shape(outputs.Part, 6, tiles(s));
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(cs::range<1>{tiles(inputs.U.shape()) * kWg}), cs::wg_desc(cs::range<1>{kWg}),
                                          cs::local_desc(cs::range<1>{static_cast<std::size_t>(kLocalElems)}));
    }

private:
    static std::size_t tiles(const cs::range<3>& s) { return (s[2] / kTX) * (s[1] / TY); }
};

// =============================================================================
// NscInit: resets the atomic accumulators (one group)
// =============================================================================
class NscInit : public cs::Component<NscInit> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::atomic_f32, 1> Dt, Ws, Sq;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs&, const Outputs& out) {
        const std::size_t tid = it.get_local_id(0);
        if (tid == 0) {
            out.Dt[cs::id<1>{0}].store(std::numeric_limits<float>::infinity());
            out.Ws[cs::id<1>{0}].store(-std::numeric_limits<float>::infinity());
        }
        if (tid < 4) out.Sq[cs::id<1>{tid}].store(0.0f);
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_output(outputs.Dt, "Dt", 1);
        declare_output(outputs.Ws, "Ws", 1);
        declare_output(outputs.Sq, "Sq", 4);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

// =============================================================================
// The deterministic reductions of the per-tile partials (fixed order: each work-item a fixed stride of tiles, 8 loads in
// flight, folded in tile order, then the group's fixed tree); norms in FP64
// =============================================================================
/// Kind 0: the four sums of squares (four groups, group q sums row 2 + q); 1: the min dt (row 0); 2: the max wave speed.
template <int Kind>
class NscDet : public cs::Component<NscDet<Kind>> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> Part;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> Det;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::size_t  n = in.Part.shape()[1], tid = it.get_local_id(0), q = it.get_group(0);
        const float* const part = cs::raw::data(in.Part);
        const auto         grp  = it.get_group();
        // each work-item: tiles tid, tid + 256, ... taken 8 at a time (8 loads in flight), folded in that order
        constexpr int kIn = 8;
        const float* const row = part + (Kind == 0 ? (2 + q) * n : Kind == 1 ? 0 : n);
        if constexpr (Kind == 0) {
            double s = 0.0;
            for (std::size_t t0 = tid; t0 < n; t0 += kWg * kIn) {
                float v[kIn];
#pragma unroll
                for (int k = 0; k < kIn; ++k) v[k] = t0 + k * kWg < n ? row[t0 + k * kWg] : 0.0f;
#pragma unroll
                for (int k = 0; k < kIn; ++k) s += static_cast<double>(v[k]);
            }
            s = cs::reduce_over_group(grp, s, cs::plus<double>());
            if (tid == 0) out.Det[cs::id<1>{q}] = static_cast<float>(s);
        } else {
            constexpr bool lo = Kind == 1;
            float          m  = lo ? std::numeric_limits<float>::infinity() : -std::numeric_limits<float>::infinity();
            for (std::size_t t0 = tid; t0 < n; t0 += kWg * kIn) {
                float v[kIn];
#pragma unroll
                for (int k = 0; k < kIn; ++k) v[k] = t0 + k * kWg < n ? row[t0 + k * kWg] : m;
#pragma unroll
                for (int k = 0; k < kIn; ++k) m = lo ? fminf(m, v[k]) : fmaxf(m, v[k]);
            }
            if constexpr (lo)
                m = cs::reduce_over_group(grp, m, cs::minimum<float>());
            else
                m = cs::reduce_over_group(grp, m, cs::maximum<float>());
            if (tid == 0) out.Det[cs::id<1>{0}] = m;
        }
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        this->declare_input(inputs.Part, "Part", cs::unspecified_shape<2>());
        this->declare_output(outputs.Det, "Det", Kind == 0 ? 4 : 1);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(cs::range<1>{(Kind == 0 ? 4u : 1u) * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

// =============================================================================
// NscError: atomic - deterministic (one group), after the deterministic reductions
// =============================================================================
class NscError : public cs::Component<NscError> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 1> DetSq, DetDt, DetWs;
    };
    struct Outputs {
        cs::Output<DType::atomic_f32, 1> Dt, Ws, Sq;  ///< tied with the accumulators: read here
        cs::Output<DType::f32, 1>        Err;         ///< [dt, ws, sq0, sq1, sq2, sq3]
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        // in a fused group with waits the reductions are the previous segment, with no barrier in between: wait for its groups
        if (const std::size_t s = it.get_segment(); s > 0) {
            const std::size_t groups = it.get_segment_group_count(s - 1)[0];
            for (std::size_t g = 0; g < groups; ++g) it.wait_group(s - 1, cs::id<1>{g});
        }
        const std::size_t tid = it.get_local_id(0);
        if (tid == 0) {
            out.Err[cs::id<1>{0}] = out.Dt[cs::id<1>{0}].load() - in.DetDt[cs::id<1>{0}];
            out.Err[cs::id<1>{1}] = out.Ws[cs::id<1>{0}].load() - in.DetWs[cs::id<1>{0}];
        }
        if (tid < 4) out.Err[cs::id<1>{2 + tid}] = out.Sq[cs::id<1>{tid}].load() - in.DetSq[cs::id<1>{tid}];
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.DetSq, "DetSq", 4);
        declare_input(inputs.DetDt, "DetDt", 1);
        declare_input(inputs.DetWs, "DetWs", 1);
        declare_output(outputs.Dt, "Dt", 1);
        declare_output(outputs.Ws, "Ws", 1);
        declare_output(outputs.Sq, "Sq", 4);
        declare_output(outputs.Err, "Err", 6);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

CS_COMPONENTS_END

}  // namespace nsc
