// Reverse-mode derivative of an oclFluid3D rollout on gfx950 (an addition; no reference counterpart).
//
// The step being differentiated is the one compute() runs (libCL/phys/oclFluid3D.cpp:293-342):
//   rho_i = max(1, m Kc sum_j (h^2 - r_ij^2)^3)        (count_i > 2, else rho0)   cl:150-200
//   p_i   = p0 + k (rho_i - rho0)                      (count_i > 2, else p0)
//   F_i   = F_in + m [Kp sum_j (p_i+p_j)/(rho_i rho_j) (h-r)^2 d/r + Kv sum_j (v_j-v_i)/(rho_i rho_j) (h-r)]
//   walls : F_i += n (2000 d_w - 256 n.v_i) where d_w > 1e-5                        cl:315-361
//   clamp : F_i *= vmax/|F_i| where |F_i| > vmax; v' = v + F dt; x' = x + v' dt     cl:258-292
// The sort permutation and the recorded neighbour visits (duplicates and stale-range visits included)
// are replayed as fixed data, so the derivative is that of the computation the forward performed.
// The rollout is checkpointed every K steps; each segment is recomputed (bit-identical to the forward:
// same per-particle arithmetic, sph_device.hpp) into a tape, then reversed step by step.
//
// Pair terms: each particle gathers the derivative of its own force with respect to its own state and,
// for symmetric (clean) pairs, the cross terms of its neighbours' forces, recomputed from the same pair
// geometry; only pairs touching a repeated bucket or a stale range scatter with float atomics. The
// parameter gradients (viscosity, stiffness, rest density) are reduced in FP64.
#include "adjoint.hpp"
#include "sph_device.hpp"

#include <cstring>

namespace mfl {
namespace {

using namespace sph;

constexpr uint32_t kFlagOverflow = 1u;

// FP64 accumulation of per-lane contributions: a wave sum, then one atomic per wave.
// Parameter gradients are summed in FP64 without atomics: every block adds its sum to its own slot
// partials[block * kGradSlots + which] (launches on one stream are ordered, so the slots need no
// atomics), and k_adj_reduce_grads adds the slots up in a fixed order at the end of backward. The sums
// are therefore reproducible run to run.
constexpr uint32_t kGradSlots = 4;

__device__ inline void block_add_f64(double v, double* __restrict__ partials, uint32_t which)
{
    __shared__ double red[kThreads / kWave];
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (uint32_t w = 0; w < blockDim.x / kWave; w++) sum += red[w];
        partials[(size_t)blockIdx.x * kGradSlots + which] += sum;
    }
    __syncthreads();  // red is reused by the next call
}

// grads[k] = sum over blocks of partials[block * kGradSlots + k], one wave per slot, fixed order.
__global__ __launch_bounds__(kThreads) void k_adj_reduce_grads(const double* __restrict__ partials, uint32_t nblocks,
                                                               double* __restrict__ grads)
{
    const uint32_t which = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (which >= kGradSlots) return;
    double v = 0.0;
    for (uint32_t b = lane; b < nblocks; b += kWave) v += partials[(size_t)b * kGradSlots + which];
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (lane == 0) grads[which] = v;
}

// The gathering reverse kernels split a particle's recorded visits over kAdjLanes lanes (visit k goes
// to lane k % kAdjLanes) and add the lanes' sums in a fixed order.
constexpr uint32_t kAdjLanes = 4;

// blocks of the widest kernel that adds parameter-gradient partials (k_adj_forces)
inline uint32_t grad_blocks(size_t n) { return grid_for(n * kAdjLanes).x; }

__device__ inline float lanes_sum(float v)
{
    v += __shfl_xor(v, 1, kWave);
    v += __shfl_xor(v, 2, kWave);
    return v;
}

__device__ inline void atomic_add3(float4* p, float x, float y, float z)
{
    float* q = reinterpret_cast<float*>(p);
    unsafeAtomicAdd(q + 0, x);
    unsafeAtomicAdd(q + 1, y);
    unsafeAtomicAdd(q + 2, z);
}

// ---- tape recording (forward recompute) ---------------------------------------------------------

// A particle's neighbourhood is clean when its 27 buckets are distinct and none holds a stale range
// (a bucket not written this step: its start is not the first particle of that key). Then every
// neighbour within h is visited exactly once, and a pair of clean particles see each other exactly
// once each: the pair is symmetric.
__device__ inline bool neighbourhood_clean(const float4& pi, float h, const uint32_t* __restrict__ cell_start,
                                           const uint32_t* __restrict__ cell, uint32_t n)
{
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    uint32_t b[27];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 27; k++) {
        b[k] = cell_hash(gx + (k % 3) - 1, gy + (k / 3) % 3 - 1, gz + k / 9 - 1);
        const uint32_t st = cell_start[b[k]];
        if (st != kEmpty && (st >= n || (cell[st] & (kBuckets - 1u)) != b[k])) ok = false;
    }
#pragma unroll
    for (int a = 1; a < 27; a++)
#pragma unroll
        for (int c = 0; c < a; c++) ok = ok && (b[a] != b[c]);
    return ok;
}

// neighbourhood_clean for a whole wave whose particles share cell (gx, gy, gz): lanes 0..26 take one
// bucket each, check it for a stale range and compare it with the other 26.
__device__ inline bool cell_clean(int gx, int gy, int gz, const uint32_t* __restrict__ cell_start,
                                  const uint32_t* __restrict__ cell, uint32_t n)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t k = lane < 27 ? lane : 0u;
    const uint32_t b = cell_hash(gx + (int)(k % 3) - 1, gy + (int)((k / 3) % 3) - 1, gz + (int)(k / 9) - 1);
    bool bad = false;
    if (lane < 27) {
        const uint32_t st = cell_start[b];
        bad = st != kEmpty && (st >= n || (cell[st] & (kBuckets - 1u)) != b);
    }
#pragma unroll
    for (uint32_t d = 1; d < 27; d++) {
        const uint32_t other = (uint32_t)__shfl((int)b, (int)((k + d) % 27), kWave);
        bad = bad || (lane < 27 && other == b);
    }
    return __ballot(bad) == 0;
}

// k_density_grp in hash order with a caller-chosen list capacity; overflow is flagged. Also records
// whether each particle's neighbourhood is clean.
__global__ __launch_bounds__(kThreads) void k_rec_density(float4* __restrict__ spos, float4* __restrict__ svel,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ cell_end, uint32_t n,
                                                          mfl_params p, uint32_t* __restrict__ nbr,
                                                          uint32_t* __restrict__ ncount, uint32_t cap,
                                                          uint32_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ cell, uint8_t* __restrict__ clean)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t grp = lane / kGroupLanes, slot = lane % kGroupLanes;
    const uint32_t wave = (blockIdx.x * kThreads + threadIdx.x) / kWave;
    const uint32_t i = wave * (kWave / kGroupLanes) + grp;
    const bool live = i < n;
    const float h = p.cellSize;
    const float h2 = h * h;
    const float h2_lo = h2 * 0.9999996f, h2_hi = h2 * 1.0000005f;
    const float4 pi = live ? spos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int gx = grid_coord(pi.x, h), gy = grid_coord(pi.y, h), gz = grid_coord(pi.z, h);
    const uint64_t live_mask = __ballot(live);
    if (live_mask == 0) return;
    const int src = __ffsll((unsigned long long)live_mask) - 1;
    const int gx0 = __shfl(gx, src, kWave), gy0 = __shfl(gy, src, kWave), gz0 = __shfl(gz, src, kWave);
    const bool same = !live || (gx == gx0 && gy == gy0 && gz == gz0);
    float acc = 0.f;
    uint32_t count = 0;
    const uint32_t grp_shift = grp * kGroupLanes;
    bool is_clean;
    if (__ballot(!same) == 0) {
        const int cx = __builtin_amdgcn_readfirstlane(gx0), cy = __builtin_amdgcn_readfirstlane(gy0),
                  cz = __builtin_amdgcn_readfirstlane(gz0);
        density_visits(cx, cy, cz, pi, i, i, slot, grp_shift, live, spos, cell_start, cell_end, n, h, h2, h2_lo, h2_hi,
                       nbr, n, acc, count, cap);
        is_clean = cell_clean(cx, cy, cz, cell_start, cell, n);
    } else {
        density_visits(gx, gy, gz, pi, i, i, slot, grp_shift, live, spos, cell_start, cell_end, n, h, h2, h2_lo, h2_hi,
                       nbr, n, acc, count, cap);
        is_clean = live && slot == 0 && neighbourhood_clean(pi, h, cell_start, cell, n);
    }
    acc += __shfl_xor(acc, 1, kWave);
    acc += __shfl_xor(acc, 2, kWave);
    acc += __shfl_xor(acc, 4, kWave);
    if (!live || slot != 0) return;
    float rho, prs;
    if (count > 2) {
        rho = fmaxf(1.0f, p.mass * p.kernelConstant * acc);
        prs = p.pressure + p.stiffness * (rho - p.density);
    } else {
        rho = p.density;
        prs = p.pressure;
    }
    spos[i].w = rho;
    svel[i].w = prs;
    ncount[i] = count;
    clean[i] = is_clean ? 1 : 0;
    if (count > cap) atomicOr(flags, kFlagOverflow);
}

// k_forces_step in hash order, also keeping the force before the velocity clamp (fpre, sorted order).
__global__ __launch_bounds__(kThreads) void k_rec_forces_step(const float4* __restrict__ spos,
                                                              const float4* __restrict__ svel, float4* __restrict__ pos,
                                                              float4* __restrict__ vel, float4* __restrict__ force,
                                                              const uint32_t* __restrict__ index,
                                                              const uint32_t* __restrict__ cell_start,
                                                              const uint32_t* __restrict__ cell_end,
                                                              const uint32_t* __restrict__ nbr,
                                                              const uint32_t* __restrict__ ncount, uint32_t cap,
                                                              uint32_t n, mfl_params p, float4 box_min, float4 box_max,
                                                              float4 gravity_force, float4* __restrict__ fpre,
                                                              uint32_t* __restrict__ tape_index,
                                                              float4* __restrict__ tape_spos,
                                                              float4* __restrict__ tape_svel)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float4 pi = spos[i], vi = svel[i];
    tape_spos[i] = pi;  // sorted state with density and pressure: the tape of this step
    tape_svel[i] = vi;
    float3 P, V;
    const uint32_t count = ncount[i];
    if (count <= cap)
        sph_force_sum_nbr<2>(spos, svel, nbr, n, count, i, pi, vi, p.cellSize, P, V);
    else
        sph_force_sum(spos, svel, cell_start, cell_end, i, n, pi, vi, p.cellSize, P, V);
    const uint32_t o = index[i];
    tape_index[i] = o;
    if (o >= n) return;
    float4 f = add_sph_force(force[o], P, V, p);
    f = clip_box(f, pi, vi, p.particleRadius, box_min, box_max);
    fpre[i] = f;
    const float4 v = integrate_force(vel[o], f, p);
    force[o] = gravity_force;
    vel[o] = v;
    pos[o] = f4_add(pos[o], f4_scale(v, p.deltaTime));
}

// ---- reverse step ---------------------------------------------------------------------------------

// Euler, clamp and walls: from (xbar, vbar) of the step's outputs to the force adjoint fbar and the
// direct parts of the input adjoints, all in sorted order.
__global__ __launch_bounds__(kThreads) void k_adj_seed(const float4* __restrict__ spos, const float4* __restrict__ svel,
                                                       const uint32_t* __restrict__ index,
                                                       const float4* __restrict__ fpre,
                                                       const float4* __restrict__ xbar, const float4* __restrict__ vbar,
                                                       uint32_t n, mfl_params p, float4 mn, float4 mx,
                                                       float4* __restrict__ xb_s, float4* __restrict__ vb_s,
                                                       float2* __restrict__ rp_s, float4* __restrict__ fbar_s)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = index[i];
    if (o >= n) return;
    const float4 xb = xbar[o], vb = vbar[o];
    const float dt = p.deltaTime;
    // x' = x + v' dt, v' = v + F_c dt
    const float3 vt = make_float3(vb.x + dt * xb.x, vb.y + dt * xb.y, vb.z + dt * xb.z);
    float3 fb = make_float3(dt * vt.x, dt * vt.y, dt * vt.z);
    // clamp F_c = F vmax/|F| (cl:264-268): adjoint (vmax/|F|)(I - F^ F^T)
    const float4 F = fpre[i];
    const float speed = sqrtf(F.x * F.x + F.y * F.y + F.z * F.z + F.w * F.w);
    if (speed > p.velocitylimit) {
        const float inv = 1.0f / speed;
        const float hx = F.x * inv, hy = F.y * inv, hz = F.z * inv;
        const float dotp = hx * fb.x + hy * fb.y + hz * fb.z;
        const float sc = p.velocitylimit * inv;
        fb = make_float3(sc * (fb.x - hx * dotp), sc * (fb.y - hy * dotp), sc * (fb.z - hz * dotp));
    }
    // walls (cl:324-357): F += n (2000 d - 256 n.v), d = r - n.(x - wall)
    const float4 x = spos[i];
    const float r = p.particleRadius;
    constexpr float eps = 0.00001f;
    float3 xw = make_float3(0.f, 0.f, 0.f), vw = make_float3(0.f, 0.f, 0.f);
    // y-, y+, z-, z+, x-, x+ : component, sign of n, wall distance
    const float dist[6] = {x.y - mn.y, mx.y - x.y, x.z - mn.z, mx.z - x.z, x.x - mn.x, mx.x - x.x};
#pragma unroll
    for (int w = 0; w < 6; w++) {
        if (r - dist[w] > eps) {
            const int axis = w < 2 ? 1 : (w < 4 ? 2 : 0);
            const float fn = (axis == 0 ? fb.x : (axis == 1 ? fb.y : fb.z));  // n.fbar up to the sign of n
            // d(n s)/dx = n (2000 dd/dx) with dd/dx = -n: -2000 n n^T ; d(n s)/dv = -256 n n^T
            if (axis == 0) {
                xw.x -= 2000.0f * fn;
                vw.x -= 256.0f * fn;
            } else if (axis == 1) {
                xw.y -= 2000.0f * fn;
                vw.y -= 256.0f * fn;
            } else {
                xw.z -= 2000.0f * fn;
                vw.z -= 256.0f * fn;
            }
        }
    }
    xb_s[i] = make_float4(xb.x + xw.x, xb.y + xw.y, xb.z + xw.z, 0.f);
    vb_s[i] = make_float4(vt.x + vw.x, vt.y + vw.y, vt.z + vw.z, 0.f);
    rp_s[i] = make_float2(0.f, 0.f);
    fbar_s[i] = make_float4(fb.x, fb.y, fb.z, 0.f);
}

// Pressure + viscosity pair terms of one recorded visit (home a, neighbour b, d = x_a - x_b) and
// their adjoint for the home's force cotangent g = m fbar_a: the derivative with respect to d
// (x_a gets +dbar, x_b gets -dbar), v_b (+vb; v_a gets -vb), rho_a, rho_b, p (both get pb), and the
// viscosity sum U that dL/dKv needs.
struct PairAdj {
    float dbx, dby, dbz, vbx, vby, vbz, rb_a, rb_b, pb, ux, uy, uz;
};

__device__ inline PairAdj pair_force_adjoint(const float4& pa, const float4& va, const float4& pb_, const float4& vb_,
                                             float gx, float gy, float gz, float kp, float kv, float h)
{
    PairAdj o;
    const float dx = pa.x - pb_.x, dy = pa.y - pb_.y, dz = pa.z - pb_.z;
    const float r2 = dist2(dx, dy, dz);
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float rr = r2 * inv_r;
    const float hr = h - rr;
    const float inv_d2 = __builtin_amdgcn_rcpf(pa.w * pb_.w);
    const float q = (va.w + vb_.w) * inv_d2;
    const float s = q * hr * hr * inv_r;
    // pressure term s d: dbar = s g + (g.d) ds/dr d / r with ds/dr = -q hr (h + r) / r^2
    const float gpx = kp * gx, gpy = kp * gy, gpz = kp * gz;
    const float gd = gpx * dx + gpy * dy + gpz * dz;
    const float dsdr_r = -q * hr * (h + rr) * inv_r * inv_r * inv_r;
    o.dbx = s * gpx + gd * dsdr_r * dx;
    o.dby = s * gpy + gd * dsdr_r * dy;
    o.dbz = s * gpz + gd * dsdr_r * dz;
    const float qbar = gd * hr * hr * inv_r;
    o.pb = qbar * inv_d2;
    const float inv_rho_a = 1.0f / pa.w, inv_rho_b = 1.0f / pb_.w;
    o.rb_a = -qbar * q * inv_rho_a;
    o.rb_b = -qbar * q * inv_rho_b;
    // viscosity term w (v_b - v_a), w = (h - r) / (rho_a rho_b): dw/dr = -1 / (rho_a rho_b)
    const float gvx = kv * gx, gvy = kv * gy, gvz = kv * gz;
    const float dvx = vb_.x - va.x, dvy = vb_.y - va.y, dvz = vb_.z - va.z;
    const float w = hr * inv_d2;
    const float gu = gvx * dvx + gvy * dvy + gvz * dvz;
    const float c = -gu * inv_d2 * inv_r;
    o.dbx += c * dx;
    o.dby += c * dy;
    o.dbz += c * dz;
    o.rb_a -= w * gu * inv_rho_a;
    o.rb_b -= w * gu * inv_rho_b;
    o.vbx = w * gvx;
    o.vby = w * gvy;
    o.vbz = w * gvz;
    o.ux = w * dvx;
    o.uy = w * dvy;
    o.uz = w * dvz;
    return o;
}

// Adjoint of the pressure + viscosity sums (sorted order), gathered per particle j over its recorded
// visits i: j's own terms (the derivative of F_j), and, when both j and i are clean, i's cross terms
// (the derivative of F_i with respect to j's state), which j reproduces from the same pair geometry.
// A pair touching an unclean particle is not symmetric (repeated buckets, stale ranges), so there each
// visit scatters its cross terms to the neighbour with atomics instead.
__global__ __launch_bounds__(kThreads) void k_adj_forces(const float4* __restrict__ spos,
                                                         const float4* __restrict__ svel,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ ncount,
                                                         const uint8_t* __restrict__ clean, uint32_t n, mfl_params p,
                                                         const float4* __restrict__ fbar_s, float4* __restrict__ xb_s,
                                                         float4* __restrict__ vb_s, float2* __restrict__ rp_s,
                                                         double* __restrict__ grads)
{
    const uint32_t j = (blockIdx.x * kThreads + threadIdx.x) / kAdjLanes, sub = threadIdx.x % kAdjLanes;
    double kv_grad = 0.0;
    if (j < n) {
        const float h = p.cellSize;
        const float kp = p.pressureConstant, kv = p.viscosityConstant;
        const float4 pj = spos[j], vj = svel[j];
        const float4 fj = fbar_s[j];
        const float gjx = p.mass * fj.x, gjy = p.mass * fj.y, gjz = p.mass * fj.z;
        const bool clean_j = clean[j] != 0;
        float xbx = 0.f, xby = 0.f, xbz = 0.f, vbx = 0.f, vby = 0.f, vbz = 0.f, rb = 0.f, pbs = 0.f;
        float ux = 0.f, uy = 0.f, uz = 0.f;
        const uint32_t count = ncount[j];
        for (uint32_t k = sub; k < count; k += kAdjLanes) {
            const uint32_t i = nbr[nbr_at(j, k, n)];
            const float4 pi = spos[i], vi = svel[i];
            // own: F_j's pair with i
            const PairAdj a = pair_force_adjoint(pj, vj, pi, vi, gjx, gjy, gjz, kp, kv, h);
            xbx += a.dbx;
            xby += a.dby;
            xbz += a.dbz;
            vbx -= a.vbx;
            vby -= a.vby;
            vbz -= a.vbz;
            rb += a.rb_a;
            pbs += a.pb;
            ux += a.ux;
            uy += a.uy;
            uz += a.uz;
            if (clean_j && clean[i] != 0) {
                // cross: F_i's pair with j, seen from j
                const float4 fi = fbar_s[i];
                const PairAdj c = pair_force_adjoint(pi, vi, pj, vj, p.mass * fi.x, p.mass * fi.y, p.mass * fi.z, kp,
                                                     kv, h);
                xbx -= c.dbx;
                xby -= c.dby;
                xbz -= c.dbz;
                vbx += c.vbx;
                vby += c.vby;
                vbz += c.vbz;
                rb += c.rb_b;
                pbs += c.pb;
            } else {
                atomic_add3(&xb_s[i], -a.dbx, -a.dby, -a.dbz);
                atomic_add3(&vb_s[i], a.vbx, a.vby, a.vbz);
                float* rpi = reinterpret_cast<float*>(&rp_s[i]);
                unsafeAtomicAdd(rpi + 0, a.rb_b);
                unsafeAtomicAdd(rpi + 1, a.pb);
            }
        }
        xbx = lanes_sum(xbx);
        xby = lanes_sum(xby);
        xbz = lanes_sum(xbz);
        vbx = lanes_sum(vbx);
        vby = lanes_sum(vby);
        vbz = lanes_sum(vbz);
        rb = lanes_sum(rb);
        pbs = lanes_sum(pbs);
        ux = lanes_sum(ux);
        uy = lanes_sum(uy);
        uz = lanes_sum(uz);
        if (sub == 0) {
            atomic_add3(&xb_s[j], xbx, xby, xbz);
            atomic_add3(&vb_s[j], vbx, vby, vbz);
            float* rpj = reinterpret_cast<float*>(&rp_s[j]);
            unsafeAtomicAdd(rpj + 0, rb);
            unsafeAtomicAdd(rpj + 1, pbs);
            // dL/dKv = sum_j m fbar_j . V_j
            kv_grad = (double)p.mass * ((double)fj.x * ux + (double)fj.y * uy + (double)fj.z * uz);
        }
    }
    block_add_f64(kv_grad, grads, 0);
}

// Adjoint of p = p0 + k (rho - rho0) and rho = max(1, m Kc acc) (count > 2), or rho = rho0: the
// cotangent of acc and the stiffness / rest-density gradients.
__global__ __launch_bounds__(kThreads) void k_adj_pressure(const float4* __restrict__ spos,
                                                           const uint32_t* __restrict__ ncount, uint32_t n,
                                                           mfl_params p, const float2* __restrict__ rp_s,
                                                           float* __restrict__ accbar, double* __restrict__ grads)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double k_grad = 0.0, rho0_grad = 0.0;
    if (i < n) {
        const float rho = spos[i].w;
        const float2 rp = rp_s[i];
        float rb = rp.x;
        const float pb = rp.y;
        float ab = 0.f;
        if (ncount[i] > 2) {
            rb += p.stiffness * pb;
            k_grad = (double)pb * ((double)rho - (double)p.density);
            rho0_grad = -(double)p.stiffness * pb;
            if (rho > 1.0f) ab = rb * p.mass * p.kernelConstant;
        } else {
            rho0_grad = (double)rb;  // rho = rho0 (the pressure is p0, independent of rho0)
        }
        accbar[i] = ab;
    }
    block_add_f64(k_grad, grads, 1);
    block_add_f64(rho0_grad, grads, 2);
}

// Adjoint of acc_a = sum_b (h^2 - r_ab^2)^3 (sorted order): with c_a = -6 accbar_a (h^2 - r^2)^2 the
// pair gives x_a += c_a d, x_b -= c_a d; a clean pair is gathered from both sides as (c_j + c_i) d_ji.
__global__ __launch_bounds__(kThreads) void k_adj_density(const float4* __restrict__ spos,
                                                          const uint32_t* __restrict__ nbr,
                                                          const uint32_t* __restrict__ ncount,
                                                          const uint8_t* __restrict__ clean, uint32_t n, mfl_params p,
                                                          const float* __restrict__ accbar, float4* __restrict__ xb_s)
{
    const uint32_t j = (blockIdx.x * kThreads + threadIdx.x) / kAdjLanes, sub = threadIdx.x % kAdjLanes;
    if (j >= n) return;  // whole groups of kAdjLanes lanes leave together
    const float h2 = p.cellSize * p.cellSize;
    const float4 pj = spos[j];
    const float aj = accbar[j];
    const bool clean_j = clean[j] != 0;
    float xbx = 0.f, xby = 0.f, xbz = 0.f;
    const uint32_t count = ncount[j];
    for (uint32_t k = sub; k < count; k += kAdjLanes) {
        const uint32_t i = nbr[nbr_at(j, k, n)];
        const float4 pi = spos[i];
        const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const float w = h2 - (dist2(dx, dy, dz));
        const float w2 = -6.0f * w * w;
        float c = aj * w2;
        if (clean_j && clean[i] != 0) {
            c += accbar[i] * w2;
        } else if (aj != 0.f) {
            atomic_add3(&xb_s[i], -c * dx, -c * dy, -c * dz);
        }
        xbx += c * dx;
        xby += c * dy;
        xbz += c * dz;
    }
    xbx = lanes_sum(xbx);
    xby = lanes_sum(xby);
    xbz = lanes_sum(xbz);
    if (sub == 0) atomic_add3(&xb_s[j], xbx, xby, xbz);
}

__global__ __launch_bounds__(kThreads) void k_adj_scatter(const uint32_t* __restrict__ index,
                                                          const float4* __restrict__ xb_s,
                                                          const float4* __restrict__ vb_s, uint32_t n,
                                                          float4* __restrict__ xbar, float4* __restrict__ vbar)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = index[i];
    if (o >= n) return;
    xbar[o] = xb_s[i];
    vbar[o] = vb_s[i];
}

}  // namespace

Adjoint::Adjoint(Fluid3D& fluid, uint32_t steps, uint32_t checkpoint_every, uint32_t list_cap) : f_(fluid)
{
    if (steps == 0 || checkpoint_every == 0 || list_cap == 0) {
        set_error("adjoint needs steps, checkpoint interval and list capacity > 0");
        return;
    }
    steps_ = steps;
    every_ = checkpoint_every;
    cap_ = (list_cap + 3u) & ~3u;  // lists are stored in chunks of four entries
    n_ = fluid.particle_count();
    const size_t ncp = (steps_ + every_ - 1) / every_ + 1;
    const size_t seg = every_ < steps_ ? every_ : steps_;
    const size_t v4 = n_ * 16, u1 = n_ * 4;
    ckpt_.resize(ncp);
    for (auto& c : ckpt_)
        if (!c.allocate(3 * v4 + 2 * (size_t)kBuckets * 4)) return;
    if (!t_index_.allocate(seg * u1) || !t_spos_.allocate(seg * v4) || !t_svel_.allocate(seg * v4) ||
        !t_ncount_.allocate(seg * u1) || !t_nbr_.allocate(seg * cap_ * u1) || !t_fpre_.allocate(seg * v4) ||
        !t_clean_.allocate(seg * n_) || !accbar_.allocate(u1) ||
        !xbar_.allocate(v4) || !vbar_.allocate(v4) || !xb_s_.allocate(v4) || !vb_s_.allocate(v4) ||
        !rp_s_.allocate(n_ * 8) || !fbar_s_.allocate(v4) || !grads_.allocate((grad_blocks(n_) + 1) * kGradSlots * sizeof(double)) ||
        !flags_.allocate(4))
        return;
    ok_ = true;
}

Adjoint::~Adjoint()
{
    for (auto& c : ckpt_) c.release();
    for (DeviceBuffer* b : {&t_index_, &t_spos_, &t_svel_, &t_ncount_, &t_nbr_, &t_fpre_, &t_clean_, &xbar_, &vbar_,
                            &xb_s_, &vb_s_, &rp_s_, &fbar_s_, &accbar_, &grads_, &flags_})
        b->release();
}

int Adjoint::save_checkpoint(uint32_t c, hipStream_t s)
{
    char* dst = static_cast<char*>(ckpt_[c].ptr);
    const size_t v4 = n_ * 16, tb = (size_t)kBuckets * 4;
    const int kinds[5] = {MFL_POSITION, MFL_VELOCITY, MFL_FORCE, MFL_CELL_START, MFL_CELL_END};
    const size_t sizes[5] = {v4, v4, v4, tb, tb};
    for (int k = 0; k < 5; k++) {
        DeviceBuffer* b = f_.select(kinds[k]);
        MFL_HIP(hipMemcpyAsync(dst, b->ptr, sizes[k], hipMemcpyDeviceToDevice, s));
        dst += sizes[k];
    }
    return 1;
}

int Adjoint::restore_checkpoint(uint32_t c, hipStream_t s)
{
    const char* src = static_cast<const char*>(ckpt_[c].ptr);
    const size_t v4 = n_ * 16, tb = (size_t)kBuckets * 4;
    const int kinds[5] = {MFL_POSITION, MFL_VELOCITY, MFL_FORCE, MFL_CELL_START, MFL_CELL_END};
    const size_t sizes[5] = {v4, v4, v4, tb, tb};
    for (int k = 0; k < 5; k++) {
        DeviceBuffer* b = f_.select(kinds[k]);
        MFL_HIP(hipMemcpyAsync(b->ptr, src, sizes[k], hipMemcpyDeviceToDevice, s));
        src += sizes[k];
    }
    return 1;
}

int Adjoint::forward(hipStream_t s)
{
    if (!ok_) {
        set_error("adjoint was not created");
        return 0;
    }
    // The forward runs the program's own step (any launch shape); backward recomputes each segment from
    // its checkpoint with the recording kernels, whose sums may differ from the forward's in FP32
    // summation order only, and differentiates that recomputed trajectory.
    if (f_.particle_count() != n_ || f_.integrate_cb) {
        set_error("adjoint: the program changed size or has an integrate callback");
        return 0;
    }
    if (!f_.check_user_buffers()) return 0;
    for (uint32_t step = 0; step < steps_; step++) {
        if (step % every_ == 0 && !save_checkpoint(step / every_, s)) return 0;
        if (!f_.compute(s)) return 0;
    }
    if (!save_checkpoint((uint32_t)ckpt_.size() - 1, s)) return 0;
    have_forward_ = true;
    return 1;
}

int Adjoint::record_step(uint32_t slot, hipStream_t s)
{
    const uint32_t n = (uint32_t)n_;
    const dim3 b(kThreads), g(grid_for(n));
    if (!f_.enqueue_front(s)) return 0;
    auto* spos = f_.select(MFL_SORTED_POSITION)->as<float4>();
    auto* svel = f_.select(MFL_SORTED_VELOCITY)->as<float4>();
    auto* cs = f_.select(MFL_CELL_START)->as<uint32_t>();
    auto* ce = f_.select(MFL_CELL_END)->as<uint32_t>();
    auto* index = f_.select(MFL_INDEX)->as<uint32_t>();
    uint32_t* nbr = t_nbr_.as<uint32_t>() + (size_t)slot * cap_ * n_;
    uint32_t* ncount = t_ncount_.as<uint32_t>() + (size_t)slot * n_;
    float4* fpre = t_fpre_.as<float4>() + (size_t)slot * n_;
    hipLaunchKernelGGL(k_rec_density, grid_for((size_t)n * kGroupLanes), b, 0, s, spos, svel, cs, ce, n, f_.params, nbr,
                       ncount, cap_, flags_.as<uint32_t>(), f_.select(MFL_CELL)->as<uint32_t>(),
                       t_clean_.as<uint8_t>() + (size_t)slot * n_);
    hipLaunchKernelGGL(k_rec_forces_step, g, b, 0, s, spos, svel, f_.select(MFL_POSITION)->as<float4>(),
                       f_.select(MFL_VELOCITY)->as<float4>(), f_.select(MFL_FORCE)->as<float4>(), index, cs, ce, nbr,
                       ncount, cap_, n, f_.params, to_f4(f_.box_min), to_f4(f_.box_max), gravity_force_of(f_.gravity),
                       fpre, t_index_.as<uint32_t>() + (size_t)slot * n_, t_spos_.as<float4>() + (size_t)slot * n_,
                       t_svel_.as<float4>() + (size_t)slot * n_);
    MFL_HIP(hipGetLastError());
    return 1;
}

int Adjoint::reverse_step(uint32_t slot, hipStream_t s)
{
    const uint32_t n = (uint32_t)n_;
    const dim3 b(kThreads), g(grid_for(n));
    const uint32_t* index = t_index_.as<uint32_t>() + (size_t)slot * n_;
    const float4* spos = t_spos_.as<float4>() + (size_t)slot * n_;
    const float4* svel = t_svel_.as<float4>() + (size_t)slot * n_;
    const uint32_t* nbr = t_nbr_.as<uint32_t>() + (size_t)slot * cap_ * n_;
    const uint32_t* ncount = t_ncount_.as<uint32_t>() + (size_t)slot * n_;
    const float4* fpre = t_fpre_.as<float4>() + (size_t)slot * n_;
    hipLaunchKernelGGL(k_adj_seed, g, b, 0, s, spos, svel, index, fpre, xbar_.as<float4>(), vbar_.as<float4>(), n,
                       f_.params, to_f4(f_.box_min), to_f4(f_.box_max), xb_s_.as<float4>(), vb_s_.as<float4>(),
                       rp_s_.as<float2>(), fbar_s_.as<float4>());
    const uint8_t* clean = t_clean_.as<uint8_t>() + (size_t)slot * n_;
    const dim3 ga(grid_for((size_t)n * kAdjLanes));
    hipLaunchKernelGGL(k_adj_forces, ga, b, 0, s, spos, svel, nbr, ncount, clean, n, f_.params, fbar_s_.as<float4>(),
                       xb_s_.as<float4>(), vb_s_.as<float4>(), rp_s_.as<float2>(), grads_.as<double>());
    hipLaunchKernelGGL(k_adj_pressure, g, b, 0, s, spos, ncount, n, f_.params, rp_s_.as<float2>(),
                       accbar_.as<float>(), grads_.as<double>());
    hipLaunchKernelGGL(k_adj_density, ga, b, 0, s, spos, nbr, ncount, clean, n, f_.params, accbar_.as<float>(),
                       xb_s_.as<float4>());
    hipLaunchKernelGGL(k_adj_scatter, g, b, 0, s, index, xb_s_.as<float4>(), vb_s_.as<float4>(), n, xbar_.as<float4>(),
                       vbar_.as<float4>());
    MFL_HIP(hipGetLastError());
    return 1;
}

int Adjoint::backward(const float4* pos_bar, const float4* vel_bar, hipStream_t s)
{
    if (!have_forward_) {
        set_error("adjoint backward needs a forward rollout first");
        return 0;
    }
    const size_t v4 = n_ * 16;
    if (pos_bar)
        MFL_HIP(hipMemcpyAsync(xbar_.ptr, pos_bar, v4, hipMemcpyDeviceToDevice, s));
    else
        MFL_HIP(hipMemsetAsync(xbar_.ptr, 0, v4, s));
    if (vel_bar)
        MFL_HIP(hipMemcpyAsync(vbar_.ptr, vel_bar, v4, hipMemcpyDeviceToDevice, s));
    else
        MFL_HIP(hipMemsetAsync(vbar_.ptr, 0, v4, s));
    MFL_HIP(hipMemsetAsync(grads_.ptr, 0, grads_.bytes, s));
    MFL_HIP(hipMemsetAsync(flags_.ptr, 0, 4, s));
    const uint32_t nseg = (steps_ + every_ - 1) / every_;
    for (int c = (int)nseg - 1; c >= 0; c--) {
        const uint32_t first = (uint32_t)c * every_;
        const uint32_t len = std::min(every_, steps_ - first);
        if (!restore_checkpoint((uint32_t)c, s)) return 0;
        for (uint32_t l = 0; l < len; l++)
            if (!record_step(l, s)) return 0;
        for (int l = (int)len - 1; l >= 0; l--)
            if (!reverse_step((uint32_t)l, s)) return 0;
    }
    if (!restore_checkpoint((uint32_t)ckpt_.size() - 1, s)) return 0;
    const uint32_t nblocks = grad_blocks(n_);
    hipLaunchKernelGGL(k_adj_reduce_grads, dim3(1), dim3(kThreads), 0, s, grads_.as<double>(), nblocks,
                       grads_.as<double>() + (size_t)nblocks * kGradSlots);
    MFL_HIP(hipGetLastError());
    uint32_t flags = 0;
    MFL_HIP(hipMemcpyAsync(&flags, flags_.ptr, 4, hipMemcpyDeviceToHost, s));
    MFL_HIP(hipStreamSynchronize(s));
    if (flags & kFlagOverflow) {
        set_error("adjoint: a particle had more neighbour visits than the list capacity; raise list_cap");
        return 0;
    }
    return 1;
}

int Adjoint::get_buffer(int which, void** ptr, size_t* bytes)
{
    DeviceBuffer* b = nullptr;
    if (which == MFL_ADJ_POSITION_GRAD) b = &xbar_;
    if (which == MFL_ADJ_VELOCITY_GRAD) b = &vbar_;
    if (!b) {
        set_error("unknown adjoint buffer " + std::to_string(which));
        return 0;
    }
    if (ptr) *ptr = b->ptr;
    if (bytes) *bytes = b->bytes;
    return 1;
}

int Adjoint::param_grads(double* out4)
{
    double g[4];
    MFL_HIP(hipDeviceSynchronize());
    MFL_HIP(hipMemcpy(g, grads_.as<double>() + (size_t)grad_blocks(n_) * kGradSlots, sizeof(g), hipMemcpyDeviceToHost));
    // [0] accumulates dL/dKv; viscosityConstant = viscosity * 45/(pi h^6) (cl:146)
    const double kv = f_.params.viscosityConstant, mu = f_.params.viscosity;
    out4[0] = mu != 0.0 ? g[0] * kv / mu : 0.0;  // dL/d viscosity
    out4[1] = g[1];                              // dL/d stiffness
    out4[2] = g[2];                              // dL/d rest density
    out4[3] = g[0];                              // dL/d viscosityConstant
    return 1;
}

}  // namespace mfl
