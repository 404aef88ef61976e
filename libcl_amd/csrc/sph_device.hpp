// Per-particle device arithmetic of the SPH step, shared by the single-scene program (fluid3d.hip)
// and the batched scenes (batch.hip), so that both compute bit-identical results.
#pragma once

#include "mfl_common.hpp"

namespace mfl {
namespace sph {

constexpr int kThreads = 256;

// XCD-aware block order. Workgroups are dealt to the 8 XCDs round robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2. Mapping b to chunk (b % 8) * (grid / 8) + b / 8 hands every XCD
// one contiguous run of the traversal order, so particles that share neighbours share an L2. Grids
// launched with it are a multiple of 8 blocks (grid_xcd).

__device__ inline uint32_t xcd_block()
{
    const uint32_t per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

inline dim3 grid_xcd(size_t threads)
{
    const size_t blocks = (threads + kThreads - 1) / kThreads;
    return dim3((unsigned)((blocks + kXcds - 1) / kXcds * kXcds));
}

__device__ inline float4 f4_add(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

__device__ inline float4 f4_scale(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }

// Distance test of cl:177-178 / cl:232-233: vlen = sqrt(x^2+y^2+z^2) < h. A candidate whose squared
// distance exceeds h^2 by more than the sqrt rounding can affect is rejected without the sqrt.
__device__ inline bool within(float r2, float h, float h2_hi, float* vlen)
{
    if (r2 > h2_hi) return false;
    *vlen = sqrtf(r2);
    return *vlen < h;
}

// Pressure and viscosity terms of one accepted pair (cl:237-242):
//   P += (p_i + p_j) / (rho_i rho_j) * (d / |d|) * (h - |d|)^2,   V += (v_j - v_i) / (rho_i rho_j) * (h - |d|)
// with d = x_i - x_j and r2 = |d|^2. Both force paths (recorded visits and full traversal) use it, so
// their sums are bit-identical.
__device__ inline void pair_force(const float4& pi, const float4& vi, const float4& pj, const float4& vj, float dx,
                                  float dy, float dz, float r2, float h, float3& P, float3& V, bool use = true)
{
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float hr = fmaf(-r2, inv_r, h);
    const float inv_d2 = __builtin_amdgcn_rcpf(pi.w * pj.w);
    const float w = hr * inv_d2;
    const float s = (vi.w + vj.w) * w * hr * inv_r;  // (p_i + p_j) (h - r)^2 / (rho_i rho_j r), w reused
    // `use` = false computes the terms and keeps the sums (a select, not a branch)
    P.x = use ? fmaf(s, dx, P.x) : P.x;
    P.y = use ? fmaf(s, dy, P.y) : P.y;
    P.z = use ? fmaf(s, dz, P.z) : P.z;
    V.x = use ? fmaf(w, vj.x - vi.x, V.x) : V.x;
    V.y = use ? fmaf(w, vj.y - vi.y, V.y) : V.y;
    V.z = use ? fmaf(w, vj.z - vi.z, V.z) : V.z;
}

// Exact form of the reference test fast_length(d) < h (cl:177-178, cl:232-233): accepted when
// sqrt(r2) < h. r2 below h2_lo (h^2 (1 - 2^-22)) or above h2_hi (h^2 (1 + 5e-7)) is decided without
// the square root.
__device__ inline bool accept_r2(float r2, float h, float h2_lo, float h2_hi)
{
    if (r2 < h2_lo) return true;
    if (r2 > h2_hi) return false;
    return sqrtf(r2) < h;
}

// Pressure and viscosity sums of cl:213-249 for sorted particle i by the full 27-bucket traversal (the
// fallback when a particle has more visits than its recorded list holds).
// `first` and `stride` select the visits first, first + stride, ... of every range (all of them by
// default), so that the lanes of one particle can split the walk.
__device__ inline void sph_force_sum(const float4* __restrict__ spos, const float4* __restrict__ svel,
                                     const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
                                     uint32_t i, uint32_t n, const float4& pi, const float4& vi, float h, float3& P,
                                     float3& V, uint32_t first = 0, uint32_t stride = 1)
{
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    P = make_float3(0.f, 0.f, 0.f);
    V = make_float3(0.f, 0.f, 0.f);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                const uint32_t c = cell_hash(gx + x, gy + y, gz + z);
                const uint32_t start = cell_start[c];
                if (start == kEmpty) continue;
                const uint32_t end = min(cell_end[c], n);  // stale or written ranges always lie in [0, n]
                for (uint32_t j = start + first; j < end; j += stride) {
                    if (j == i) continue;
                    const float4 pj = spos[j];
                    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const float r2 = dist2(dx, dy, dz);
                    if (accept_r2(r2, h, h2_lo, h2_hi)) pair_force(pi, vi, pj, svel[j], dx, dy, dz, r2, h, P, V);
                }
            }
}

// Walks of a neighbourhood that overflowed the LDS tile (the reference's loops, cl:161-190 and
// cl:213-249), split over the `stride` lanes of a particle: lane `first` takes the visits first,
// first + stride, ... of every bucket range, so each visit is tested once and the caller adds the
// lanes' sums. Out of line (the rare path must not set the tile kernels' register budget) and returning
// by value (a pointer to the caller's sums would keep them in scratch memory).
struct WalkSums {
    float acc;
    uint32_t count;
};

__device__ __noinline__ inline WalkSums density_walk_split(int cx, int cy, int cz, float4 pi, uint32_t i,
                                                          const float4* __restrict__ spos,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t n, float h,
                                                          uint32_t first, uint32_t stride)
{
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    WalkSums r{0.f, 0u};
    for (int k = 0; k < 27; k++) {
        const uint32_t b = cell_hash(cx + k % 3 - 1, cy + (k / 3) % 3 - 1, cz + k / 9 - 1);
        const uint32_t st = cell_start[b];
        if (st == kEmpty) continue;
        const uint32_t end = min(cell_end[b], n);
        for (uint32_t j = st + first; j < end; j += stride) {
            const float4 pj = spos[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2(dx, dy, dz);
            if (j != i && accept_r2(r2, h, h2_lo, h2_hi)) {
                const float w = h2 - r2;
                r.acc = fmaf(w * w, w, r.acc);
                r.count++;
            }
        }
    }
    return r;
}

struct WalkForces {
    float3 P, V;
};

__device__ __noinline__ inline WalkForces force_walk_split(const float4* __restrict__ spos,
                                                          const float4* __restrict__ svel,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t i,
                                                          uint32_t n, float4 pi, float4 vi, float h, uint32_t first,
                                                          uint32_t stride)
{
    WalkForces r;
    sph_force_sum(spos, svel, cell_start, cell_end, i, n, pi, vi, h, r.P, r.V, first, stride);
    return r;
}

__device__ inline float4 add_sph_force(float4 f, const float3& P, const float3& V, const mfl_params& p)
{
    f.x += (p.pressureConstant * P.x + p.viscosityConstant * V.x) * p.mass;
    f.y += (p.pressureConstant * P.y + p.viscosityConstant * V.y) * p.mass;
    f.z += (p.pressureConstant * P.z + p.viscosityConstant * V.z) * p.mass;
    f.w = 0.f;  // cl:253
    return f;
}

// repulsionForce (cl:309-312) accumulated as in cl:327-357.
__device__ inline void add_repulsion(float4& f, float4 nrm, float4 vel, float dist)
{
    const float dot = nrm.x * vel.x + nrm.y * vel.y + nrm.z * vel.z + nrm.w * vel.w;
    const float s = 2000.0f * dist - 256.0f * dot;
    f = f4_add(f, f4_scale(nrm, s));
}

// clClipBox body (cl:319-358) on a private force value.
__device__ inline float4 clip_box(float4 f, float4 pos, float4 vel, float r, float4 mn, float4 mx)
{
    constexpr float eps = 0.00001f;  // cl:304
    float diff = r - (pos.y - mn.y);
    if (diff > eps) add_repulsion(f, make_float4(0.f, 1.f, 0.f, 0.f), vel, diff);
    diff = r - (mx.y - pos.y);
    if (diff > eps) add_repulsion(f, make_float4(0.f, -1.f, 0.f, 0.f), vel, diff);
    diff = r - (pos.z - mn.z);
    if (diff > eps) add_repulsion(f, make_float4(0.f, 0.f, 1.f, 0.f), vel, diff);
    diff = r - (mx.z - pos.z);
    if (diff > eps) add_repulsion(f, make_float4(0.f, 0.f, -1.f, 0.f), vel, diff);
    diff = r - (pos.x - mn.x);
    if (diff > eps) add_repulsion(f, make_float4(1.f, 0.f, 0.f, 0.f), vel, diff);
    diff = r - (mx.x - pos.x);
    if (diff > eps) add_repulsion(f, make_float4(-1.f, 0.f, 0.f, 0.f), vel, diff);
    return f;
}

// clIntegrateForce body (cl:262-270): clamp a private copy of the force, then v += f * dt.
__device__ inline float4 integrate_force(float4 vel, float4 f, const mfl_params& p)
{
    const float speed = sqrtf(f.x * f.x + f.y * f.y + f.z * f.z + f.w * f.w);
    if (speed > p.velocitylimit) f = f4_scale(f, p.velocitylimit / speed);
    return f4_add(vel, f4_scale(f, p.deltaTime));
}

constexpr uint32_t kNbrCap = 64;

// Recorded visit lists, slot-major in chunks of four: entry k of list t is element
// ((k / 4) * stride + t) * 4 + k % 4, so a lane reads four consecutive entries with one 16-byte load
// and the lanes of a wave read one contiguous run. A list of capacity cap needs ceil(cap / 4) * 4 *
// stride elements.
__device__ inline size_t nbr_at(uint32_t t, uint32_t k, size_t stride)
{
    return (((size_t)(k >> 2) * stride + t) << 2) | (k & 3u);
}

// Density pass with eight lanes per particle. A wave takes eight consecutive entries of the traversal
// order (in Morton or hash order usually the particles of one cell), and the eight lanes of a group
// split each bucket range of their particle into interleaved slots, so bucket trip counts are the
// same for the whole wave. When all eight particles share a cell, the 27 bucket hashes and range
// loads are wave-uniform (scalar). Accepted visits are appended in the reference's visit order
// (bucket by bucket, ascending j) with a group ballot.
constexpr uint32_t kGroupLanes = 8;

__device__ inline void density_visits(int gx, int gy, int gz, const float4& pi, uint32_t i, uint32_t t, uint32_t slot,
                                      uint32_t grp_shift, bool live, const float4* __restrict__ spos,
                                      const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
                                      uint32_t n, float h, float h2, float h2_lo, float h2_hi,
                                      uint32_t* __restrict__ nbr, size_t nbr_stride, float& acc, uint32_t& count,
                                      uint32_t cap = kNbrCap)
{
    const uint32_t below = (1u << slot) - 1u;
    for (int z = -1; z <= 1; z++) {
        uint32_t st[9], en[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const uint32_t c = cell_hash(gx + (k % 3) - 1, gy + (k / 3) - 1, gz + z);
            st[k] = cell_start[c];
            en[k] = cell_end[c];
        }
#pragma unroll
        for (int k = 0; k < 9; k++) {
            if (st[k] == kEmpty) continue;
            const uint32_t end = min(en[k], n);
            for (uint32_t j0 = st[k]; j0 < end; j0 += kGroupLanes) {
                const uint32_t j = j0 + slot;
                bool ok = false;
                float r2 = 0.f;
                if (j < end) {
                    const float4 pj = spos[j];
                    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    r2 = dist2(dx, dy, dz);
                    ok = live && j != i && accept_r2(r2, h, h2_lo, h2_hi);
                }
                const uint32_t mask = (uint32_t)(__ballot(ok) >> grp_shift) & 0xFFu;
                if (ok) {
                    const float w = h2 - r2;  // h^2 - vlen^2 of cl:123
                    acc = fmaf(w * w, w, acc);
                    const uint32_t at = count + (uint32_t)__popc(mask & below);
                    if (at < cap) nbr[nbr_at(t, at, nbr_stride)] = j;
                }
                count += (uint32_t)__popc(mask);
            }
        }
    }
}

// Pressure and viscosity sums of cl:237-242 over the visits recorded for thread t (sorted particle i).
// CH 16-byte chunks of the list (4 CH visits) per trip: their entries, then the positions and
// velocities of their visits, are in flight together; the sums still run in list order.
template <int CH = 1>
__device__ inline void sph_force_sum_nbr(const float4* __restrict__ spos, const float4* __restrict__ svel,
                                         const uint32_t* __restrict__ nbr, size_t nbr_stride, uint32_t count,
                                         uint32_t t, const float4& pi, const float4& vi, float h, float3& P, float3& V)
{
    constexpr uint32_t U = 4 * CH;
    P = make_float3(0.f, 0.f, 0.f);
    V = make_float3(0.f, 0.f, 0.f);
    if (count == 0) return;
    const uint32_t last = (count - 1) & ~3u;  // first entry of the last chunk
    for (uint32_t k = 0; k < count; k += U) {
        uint4 c[CH];
#pragma unroll
        for (int q = 0; q < CH; q++)
            c[q] = *reinterpret_cast<const uint4*>(nbr + nbr_at(t, min(k + 4 * q, last), nbr_stride));
        const uint32_t m = min(count - k, U);
        // past the end of the list a chunk holds no visit: load the first entry again instead, so that
        // all loads issue together
        uint32_t j[U];
#pragma unroll
        for (int q = 0; q < CH; q++) {
            j[4 * q + 0] = 4 * q + 0 < m ? c[q].x : c[0].x;
            j[4 * q + 1] = 4 * q + 1 < m ? c[q].y : c[0].x;
            j[4 * q + 2] = 4 * q + 2 < m ? c[q].z : c[0].x;
            j[4 * q + 3] = 4 * q + 3 < m ? c[q].w : c[0].x;
        }
        float4 pj[U], vj[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            pj[u] = spos[j[u]];
            vj[u] = svel[j[u]];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const float dx = pi.x - pj[u].x, dy = pi.y - pj[u].y, dz = pi.z - pj[u].z;
            pair_force(pi, vi, pj[u], vj[u], dx, dy, dz, dist2(dx, dy, dz), h, P, V, u < m);
        }
    }
}

// Morton code of each particle's cell (10 bits per axis, wrapped), for the listed traversal order.
__device__ inline uint32_t spread_bits3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

namespace {
// The sorted copy of the state (cl:66-72: spos[i] = pos[index[i]]) written from the caller's order
// through the inverse permutation inv (inv[index[i]] = i, every entry < n when index is a permutation):
// the hash scatters buckets over the sorted order, so gathering by index would fetch a whole line of
// pos / vel for each bucket's 32-byte runs; here the reads are coalesced and a bucket's particles
// land in consecutive slots.
__global__ __launch_bounds__(kThreads) void k_scatter_sorted(const uint32_t* __restrict__ inv,
                                                             const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel,
                                                             float4* __restrict__ spos, float4* __restrict__ svel,
                                                             uint32_t n)
{
    const uint32_t s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = inv[s];
    if (i < n) {
        spos[i] = pos[s];
        svel[i] = vel[s];
    }
}
}  // namespace

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

inline float4 to_f4(const float* v) { return make_float4(v[0], v[1], v[2], v[3]); }

// clGravity writes -9.8f * gravity (cl:297)
inline float4 gravity_force_of(const float* g)
{
    return make_float4(-9.8f * g[0], -9.8f * g[1], -9.8f * g[2], -9.8f * g[3]);
}

}  // namespace sph
}  // namespace mfl
