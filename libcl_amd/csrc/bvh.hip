// Triangle-mesh BVH build (libCL/geom/oclBvhTrimesh.cpp:53-162, kernels libCL/geom/oclBvhTrimesh.cl:31-297)
// and particle-vs-mesh collision (clCollideBVH, libCL/phys/oclFluid3D.cl:365-550) for gfx950.
//
// Build: global box (two-level min/max reduction) -> Morton key per triangle -> 32-bit stable radix sort
// -> node init (split bit of neighbouring keys, leaf boxes) -> links -> boxes, with no atomics and no
// cross-workgroup waits. The reference links each internal node by scanning left and right for the
// nearest larger split bit (O(T) per node) and walks every leaf from the root to unite the boxes. The
// tree it defines is the Cartesian tree of the split bits; where split bits tie, which only duplicate
// keys produce and where the reference's parallel writes race, the left neighbour counts as the larger,
// which keeps the tree well formed. Here each node finds its leaf range directly: for split bit b > 0
// the keys sharing key[i]'s bits above b form one contiguous run of the sorted keys (galloping search
// from i), for b = 0 the range starts at i itself (the tie rule); a node covering leaves [lo, hi] is the
// child of split lo-1 or split hi, whichever has the smaller split bit (ties: hi), and writes itself
// there. Boxes: a segment tree over the leaf boxes in leaf order, then every internal node unites its
// leaf range with an order-preserving query. The reference's box of a node, min(left, right) nested over
// its subtree with min(x, y) = y < x ? y : x, is the leftmost least value of the range (that min is
// associative for non-NaN values), which the query returns as well: the boxes are bit-identical.
//
// Definitions the OpenCL sources leave to the implementation (the oracle uses the same ones):
// min(x, y) = y < x ? y : x, max(x, y) = x < y ? y : x; dot(a, b) = ((a.x b.x + a.y b.y) + a.z b.z) +
// a.w b.w; normalize(v) = v / sqrt(dot(v, v)) (v when zero); fast_length(v) = sqrt(dot(v, v)); float
// to uint conversion truncates and saturates with NaN -> 0; the global box starts empty on every build.
#include "bvh.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <string>
#include <vector>

namespace mfl {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kBoxBlocks = 512;           // partial boxes of the global-box reduction
constexpr uint32_t kNoBound = 0xFFFFFFFFu;     // bound[] before the first child arrives
constexpr uint32_t kQueue = 256;               // QUEUE_SIZE, libCL/phys/oclFluid3D.cl:378

__device__ inline float cl_min(float x, float y) { return y < x ? y : x; }
__device__ inline float cl_max(float x, float y) { return x < y ? y : x; }
__device__ inline float4 min4(float4 a, float4 b)
{
    return make_float4(cl_min(a.x, b.x), cl_min(a.y, b.y), cl_min(a.z, b.z), cl_min(a.w, b.w));
}
__device__ inline float4 max4(float4 a, float4 b)
{
    return make_float4(cl_max(a.x, b.x), cl_max(a.y, b.y), cl_max(a.z, b.z), cl_max(a.w, b.w));
}
__device__ inline float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ inline float4 sub4(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ inline float4 scale4(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ inline float dot4(float4 a, float4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
__device__ inline float4 cross4(float4 a, float4 b)
{
    return make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f);
}
__device__ inline float4 normalize4(float4 v)
{
    if (v.x == 0.f && v.y == 0.f && v.z == 0.f && v.w == 0.f) return v;
    const float l = sqrtf(dot4(v, v));
    return make_float4(v.x / l, v.y / l, v.z / l, v.w / l);
}
__device__ inline uint32_t to_uint(float q)
{
    if (!(q == q) || q <= -1.f) return 0u;
    if (q >= 4294967296.f) return 0xFFFFFFFFu;
    return (uint32_t)q;
}
__device__ inline float4 empty_min() { return make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.f); }   // cl:36
__device__ inline float4 empty_max() { return make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.f); }

// code(), libCL/geom/oclBvhTrimesh.cl:78-85: the low 10 bits of x spread to every third bit
__device__ inline uint32_t morton_code(uint32_t x)
{
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// ---- global box (clAABB, cl:31-72) -----------------------------------------------------------------

__device__ inline void block_minmax(float4& mn, float4& mx, float4* smin, float4* smax)
{
    const uint32_t t = threadIdx.x;
    smin[t] = mn;
    smax[t] = mx;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) {
            smin[t] = min4(smin[t], smin[t + s]);
            smax[t] = max4(smax[t], smax[t + s]);
        }
        __syncthreads();
    }
    mn = smin[0];
    mx = smax[0];
}

__global__ __launch_bounds__(kBlock) void k_box_partial(const float4* __restrict__ vertex, uint32_t nv,
                                                       float4* __restrict__ part, uint32_t* __restrict__ info)
{
    __shared__ float4 smin[kBlock], smax[kBlock];
    if (blockIdx.x == 0 && threadIdx.x < 2) info[threadIdx.x] = threadIdx.x == 0 ? 0xFFFFFFFFu : 0u;  // root, error
    float4 mn = empty_min(), mx = empty_max();
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nv; i += gridDim.x * kBlock) {
        const float4 v = vertex[i];
        mn = min4(mn, v);
        mx = max4(mx, v);
    }
    block_minmax(mn, mx, smin, smax);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

__global__ __launch_bounds__(kBlock) void k_box_final(float4* __restrict__ part, uint32_t nparts)
{
    __shared__ float4 smin[kBlock], smax[kBlock];
    float4 mn = empty_min(), mx = empty_max();
    for (uint32_t i = threadIdx.x; i < nparts; i += kBlock) {
        mn = min4(mn, part[2 * i]);
        mx = max4(mx, part[2 * i + 1]);
    }
    block_minmax(mn, mx, smin, smax);
    if (threadIdx.x == 0) {
        part[2 * kBoxBlocks] = mn;  // the global box follows the partials
        part[2 * kBoxBlocks + 1] = mx;
    }
}

// ---- Morton keys (clMorton, cl:87-99) ----------------------------------------------------------------

__device__ inline bool triangle_vertices(const uint32_t* __restrict__ index, const float4* __restrict__ vertex,
                                         uint32_t nv, uint32_t tri, float4& v0, float4& v1, float4& v2)
{
    const uint32_t i0 = index[3 * (size_t)tri], i1 = index[3 * (size_t)tri + 1], i2 = index[3 * (size_t)tri + 2];
    const bool ok = i0 < nv && i1 < nv && i2 < nv;
    v0 = vertex[ok ? i0 : 0u];
    v1 = vertex[ok ? i1 : 0u];
    v2 = vertex[ok ? i2 : 0u];
    return ok;
}

__global__ __launch_bounds__(kBlock) void k_morton(const uint32_t* __restrict__ index, const float4* __restrict__ vertex,
                                                  uint32_t nv, uint32_t ntri, const float4* __restrict__ box,
                                                  uint32_t* __restrict__ key, uint32_t* __restrict__ info)
{
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= ntri) return;
    float4 v0, v1, v2;
    if (!triangle_vertices(index, vertex, nv, g, v0, v1, v2)) info[1] = 1u;  // index out of range
    const float4 mn = box[0], mx = box[1];
    const float4 s = add4(add4(v0, v1), v2);
    const float nx = (s.x / 3.0f - mn.x) / (mx.x - mn.x);
    const float ny = (s.y / 3.0f - mn.y) / (mx.y - mn.y);
    const float nz = (s.z / 3.0f - mn.z) / (mx.z - mn.z);
    key[g] = morton_code(to_uint(nx * 1023.0f)) | (morton_code(to_uint(ny * 1023.0f)) << 1) |
             (morton_code(to_uint(nz * 1023.0f)) << 2);
}

// ---- nodes (clCreateNodes cl:115-142, clCreateLeaves cl:198-237, leaf boxes cl:255-271) ----------------

// the highest of bits 29..0 where the keys differ, 0 if none
__device__ inline uint32_t split_bit(uint32_t k0, uint32_t k1)
{
    const uint32_t x = (k0 ^ k1) & 0x3FFFFFFFu;
    return x ? 31u - (uint32_t)__clz((int)x) : 0u;
}

__global__ __launch_bounds__(kBlock) void k_nodes(const uint32_t* __restrict__ key, const uint32_t* __restrict__ val,
                                                 const uint32_t* __restrict__ index, const float4* __restrict__ vertex,
                                                 uint32_t nv, uint32_t T, BvhNode* __restrict__ nodes,
                                                 uint32_t* __restrict__ split, float4* __restrict__ seg0)
{
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= T) return;
    if (g + 1 < T) {
        const uint32_t b = split_bit(key[g], key[g + 1]);
        BvhNode d;
        d.bbMin = empty_min();
        d.bbMax = empty_max();
        d.left = -1;
        d.right = -1;
        d.bit = b;
        d.trav = 0u;
        nodes[g] = d;
        split[g] = b;
    }
    const uint32_t tri = val[g];
    float4 v0, v1, v2;
    triangle_vertices(index, vertex, nv, tri, v0, v1, v2);
    BvhNode leaf;
    leaf.bbMin = min4(min4(min4(empty_min(), v0), v1), v2);
    leaf.bbMax = max4(max4(max4(empty_max(), v0), v1), v2);
    leaf.left = 0;
    leaf.right = 0;
    leaf.bit = tri;
    leaf.trav = 0u;
    nodes[T - 1 + g] = leaf;
    seg0[2 * (size_t)g] = leaf.bbMin;  // level 0 of the box segment tree, leaf order
    seg0[2 * (size_t)g + 1] = leaf.bbMax;
}

// ---- links (clLinkNodes cl:144-196, clCreateLeaves cl:198-237) ------------------------------------------

// first index of the run of sorted keys around `at` whose key >> shift equals key[at] >> shift (dir = -1),
// or its last index (dir = +1): galloping, then bisection
__device__ inline int32_t run_end(const uint32_t* __restrict__ key, int32_t T, int32_t at, uint32_t shift, int dir)
{
    const uint32_t want = key[at] >> shift;
    int32_t good = at, step = 1, bad;
    for (;;) {
        const int32_t probe = good + dir * step;
        if (probe < 0 || probe >= T) {
            bad = dir < 0 ? -1 : T;
            break;
        }
        if ((key[probe] >> shift) != want) {
            bad = probe;
            break;
        }
        good = probe;
        step <<= 1;
    }
    while ((bad - good) * dir > 1) {
        const int32_t mid = good + (bad - good) / 2;
        if ((key[mid] >> shift) == want)
            good = mid;
        else
            bad = mid;
    }
    return good;
}

// One thread per node (internal i < T-1, leaf T-1+g): leaf range, parent, and the child pointer in the
// parent; internal nodes keep their range for the box pass, the root reports itself in info[0].
__global__ __launch_bounds__(kBlock) void k_link(const uint32_t* __restrict__ key, const uint32_t* __restrict__ split,
                                                uint32_t T, BvhNode* __restrict__ nodes, uint2* __restrict__ range,
                                                uint32_t* __restrict__ info)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 2 * T - 1) return;
    uint32_t lo, hi;
    if (t < T - 1) {
        const uint32_t b = split[t];
        lo = b == 0u ? t : (uint32_t)run_end(key, (int32_t)T, (int32_t)t, b + 1u, -1);
        hi = (uint32_t)run_end(key, (int32_t)T, (int32_t)t + 1, b + 1u, +1);
        range[t] = make_uint2(lo, hi);
        if (lo == 0u && hi == T - 1u) {
            info[0] = t;
            return;
        }
    } else {
        lo = hi = t - (T - 1u);
    }
    const bool to_hi = lo == 0u || (hi != T - 1u && !(split[lo - 1u] < split[hi]));
    if (to_hi)
        nodes[hi].left = (int32_t)t;
    else
        nodes[lo - 1u].right = (int32_t)t;
}

// ---- boxes (clComputeAABBs cl:239-297) --------------------------------------------------------------

// Levels [k+1, k+8] of the segment tree from level k (n entries): block b reduces entries
// [256 b, 256 b + 256) in LDS; level k+j entry e unites entries 2e and 2e+1 of level k+j-1 (left first).
__global__ __launch_bounds__(kBlock) void k_seg_up(float4* __restrict__ seg, const uint64_t* __restrict__ offs,
                                                  const uint32_t* __restrict__ sizes, uint32_t k, uint32_t levels)
{
    __shared__ float4 smin[kBlock], smax[kBlock];
    const uint32_t t = threadIdx.x, e = blockIdx.x * kBlock + t;
    const uint32_t n = sizes[k];
    if (e < n) {
        smin[t] = seg[2 * (offs[k] + e)];
        smax[t] = seg[2 * (offs[k] + e) + 1];
    }
    __syncthreads();
    uint32_t width = kBlock;
    for (uint32_t j = 1; j <= levels; j++) {
        width >>= 1;
        const uint32_t nj = sizes[k + j], nprev = sizes[k + j - 1];
        const uint32_t ej = blockIdx.x * width + t;  // entry of level k+j
        float4 mn, mx;
        const bool live = t < width && ej < nj;
        if (live) {
            mn = smin[2 * t];
            mx = smax[2 * t];
            if (2 * ej + 1 < nprev) {
                mn = min4(mn, smin[2 * t + 1]);
                mx = max4(mx, smax[2 * t + 1]);
            }
        }
        __syncthreads();
        if (live) {
            smin[t] = mn;
            smax[t] = mx;
            seg[2 * (offs[k + j] + ej)] = mn;
            seg[2 * (offs[k + j] + ej) + 1] = mx;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_node_boxes(const float4* __restrict__ seg, const uint64_t* __restrict__ offs,
                                                      const uint2* __restrict__ range, uint32_t T,
                                                      BvhNode* __restrict__ nodes)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i + 1 >= T) return;
    const uint2 rg = range[i];
    uint32_t l = rg.x, r = rg.y + 1u, k = 0;
    float4 lmn = empty_min(), lmx = empty_max(), rmn = empty_min(), rmx = empty_max();
    bool hl = false, hr = false;
    while (l < r) {
        const float4* lv = seg + 2 * offs[k];
        if (l & 1u) {
            const float4 a = lv[2 * l], b = lv[2 * l + 1];
            lmn = hl ? min4(lmn, a) : a;
            lmx = hl ? max4(lmx, b) : b;
            hl = true;
            l++;
        }
        if (r & 1u) {
            r--;
            const float4 a = lv[2 * r], b = lv[2 * r + 1];
            rmn = hr ? min4(a, rmn) : a;
            rmx = hr ? max4(b, rmx) : b;
            hr = true;
        }
        l >>= 1;
        r >>= 1;
        k++;
    }
    nodes[i].bbMin = hl && hr ? min4(lmn, rmn) : (hl ? lmn : rmn);
    nodes[i].bbMax = hl && hr ? max4(lmx, rmx) : (hl ? lmx : rmx);
    nodes[i].trav = 2u;  // cl:282-286: the first of the two arrivals leaves 2
}

// ---- collision (clCollideBVH, libCL/phys/oclFluid3D.cl:365-550) ----------------------------------------

// collideTriangle, cl:380-461
__device__ inline bool collide_triangle(float4 V0, float4 V1, float4 V2, float4& pos, float4& vel, float4& force,
                                        const mfl_params& p)
{
    V0.w = 0.f;
    V1.w = 0.f;
    V2.w = 0.f;
    pos.w = 0.f;
    const float4 u = sub4(V1, V0), v = sub4(V2, V0);
    float4 n = cross4(u, v);
    if (dot4(n, n) == 0.f) return false;
    const float4 w0 = sub4(pos, V0);
    const float4 disp = scale4(vel, p.deltaTime);
    const float a = dot4(n, w0);
    if (a < 0.f) return false;                        // other side of the triangle
    const float b = dot4(n, disp);
    if ((double)fabsf(b) < 0.000001) return false;    // parallel to the plane (a double literal)
    const float r = -a / b;
    if (r < 0.f) return false;                        // moving away
    const float4 I = add4(pos, scale4(disp, r));
    const float uu = dot4(u, u), uv = dot4(u, v), vv = dot4(v, v);
    const float4 w = sub4(I, V0);
    const float wu = dot4(w, u), wv = dot4(w, v);
    const float D = uv * uv - uu * vv;
    const float s = (uv * wv - vv * wu) / D;
    if (s < 0.f || s > 1.f) return false;
    const float t = (uv * wu - uu * wv) / D;
    if (t < 0.f || (s + t) > 1.f) return false;
    n = normalize4(n);
    const float4 Vc = sub4(add4(pos, disp), V0);
    const float dist = dot4(Vc, n);
    const float depth = p.particleRadius - dist;
    if (!(depth > 0.f)) return false;
    const float4 dPos = scale4(scale4(normalize4(disp), dot4(n, disp)), depth);
    const float4 dTangent = sub4(vel, scale4(n, dot4(n, vel)));
    const float coef = p.pressureConstant / p.density;
    const float hd = p.cellSize - dist;
    const float4 pv = scale4(scale4(n, hd), hd);      // PressureVariable, cl:125
    force = add4(force, make_float4(dTangent.x / p.deltaTime, dTangent.y / p.deltaTime, dTangent.z / p.deltaTime,
                                    dTangent.w / p.deltaTime));
    force = add4(force, scale4(scale4(pv, coef), p.mass));
    pos = add4(pos, dPos);
    vel = make_float4(0.f, 0.f, 0.f, 0.f);
    return true;
}

// The particle's ray as collideAABB uses it (cl:486-496): origin, 1 / normalize(velocity) per axis and
// the acceptance limit max(radius, fast_length(velocity)). The walk changes neither the position's
// x, y, z nor the velocity before the collision that ends it, so these are computed once per walk.
struct Ray {
    float ox, oy, oz, ix, iy, iz, lim;
};

__device__ inline Ray make_ray(float4 ro, float4 rd, float radius)
{
    const float4 nd = normalize4(rd);
    return Ray{ro.x, ro.y, ro.z, 1.0f / nd.x, 1.0f / nd.y, 1.0f / nd.z, cl_max(radius, sqrtf(dot4(rd, rd)))};
}

// intersectBox (cl:464-484) + collideAABB (cl:486-496) against the box (mn, mx); only x, y, z reach the
// result
__device__ inline bool collide_aabb(const Ray& r, float mnx_, float mny_, float mnz_, float mxx_, float mxy_,
                                    float mxz_)
{
    const float bx = r.ix * (mnx_ - r.ox), by = r.iy * (mny_ - r.oy), bz = r.iz * (mnz_ - r.oz);
    const float tx = r.ix * (mxx_ - r.ox), ty = r.iy * (mxy_ - r.oy), tz = r.iz * (mxz_ - r.oz);
    const float mnx = cl_min(tx, bx), mny = cl_min(ty, by), mnz = cl_min(tz, bz);
    const float mxx = cl_max(tx, bx), mxy = cl_max(ty, by), mxz = cl_max(tz, bz);
    const float largest_tmin = cl_max(cl_max(mnx, mny), cl_max(mnx, mnz));
    const float smallest_tmax = cl_min(cl_min(mxx, mxy), cl_min(mxx, mxz));
    return smallest_tmax > largest_tmin && largest_tmin < r.lim;
}

// Walk copy of the tree, written after each build: one 64-byte record per internal node with both
// children's boxes (x, y, z of bbMin and bbMax, the values of the nodes) and their walk entries: an
// internal node's index, or leaf_tag | triangle for a leaf (the node's bit). One cache line per
// visited node instead of the node plus both children.
__global__ __launch_bounds__(kBlock) void k_wide(const BvhNode* __restrict__ nodes, uint32_t T, uint32_t leaf_tag,
                                                float4* __restrict__ wide)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i + 1 >= T) return;
    const int32_t l = nodes[i].left, r = nodes[i].right;
    float4* w = wide + 4 * (size_t)i;
    if (l < 0 || r < 0 || (uint32_t)l > 2u * (T - 1u) || (uint32_t)r > 2u * (T - 1u)) {  // incomplete tree
        w[3] = make_float4(__uint_as_float(0xFFFFFFFFu), __uint_as_float(0xFFFFFFFFu), 0.f, 0.f);
        return;
    }
    const float4 lmn = nodes[l].bbMin, lmx = nodes[l].bbMax, rmn = nodes[r].bbMin, rmx = nodes[r].bbMax;
    const uint32_t el = (uint32_t)l >= T - 1u ? (leaf_tag | nodes[l].bit) : (uint32_t)l;
    const uint32_t er = (uint32_t)r >= T - 1u ? (leaf_tag | nodes[r].bit) : (uint32_t)r;
    w[0] = make_float4(lmn.x, lmn.y, lmn.z, lmx.x);
    w[1] = make_float4(lmx.y, lmx.z, rmn.x, rmn.y);
    w[2] = make_float4(rmn.z, rmx.x, rmx.y, rmx.z);
    w[3] = make_float4(__uint_as_float(el), __uint_as_float(er), 0.f, 0.f);
}

// The walk's FIFO: positions head / tail count modulo 256 exactly as the reference's ring (a push
// that makes tail == head ends the walk, cl:515-548). LdsRing keeps kRingLds slots per lane in LDS
// (slot s of a lane at ring[s * kBlock + lane]: every lane in its own bank whatever s); a walk whose
// pending entries would exceed them stops without side effects and is redone with PrivateRing, the
// full 256 slots in private (scratch) memory. A walk has no side effect before the collision that ends
// it, so the redo from the start is the same walk.
constexpr uint32_t kRingLds = 32;  // divides 256, so slot = position % 32 survives the wrap

// Slot is uint16_t for trees below 65,536 nodes (4 KiB of LDS per wave) and uint32_t above.
template <class Slot>
struct LdsRing {
    static constexpr uint32_t cap = kRingLds;
    Slot* base;
    __device__ uint32_t get(uint32_t pos) const { return base[(pos % kRingLds) * kBlock]; }
    __device__ void put(uint32_t pos, uint32_t v) { base[(pos % kRingLds) * kBlock] = (Slot)v; }
};

struct PrivateRing {
    static constexpr uint32_t cap = kQueue;
    uint32_t q[kQueue];
    __device__ uint32_t get(uint32_t pos) const { return q[pos]; }
    __device__ void put(uint32_t pos, uint32_t v) { q[pos] = v; }
};

struct Mesh {
    const float4* wide;   // k_wide records
    uint32_t ninternal, root, leaf_tag;
    const uint32_t* tri;
    uint32_t ntri;
    const float4* vertex;
    uint32_t nv;
};

enum : int { kWalkMiss = 0, kWalkHit = 1, kWalkSpill = -1 };

// clCollideBVH's loop (cl:515-548) for one particle. An index outside the mesh ends the walk.
template <class Ring>
__device__ inline int walk(Ring& ring, const Mesh& m, const mfl_params& p, float4& P, float4& V, float4& F)
{
    const Ray ray = make_ray(P, V, p.particleRadius);
    uint32_t head = 0, tail = 1;
    ring.put(0, m.root);
    while (head != tail) {
        const uint32_t e = ring.get(head);
        head = (head + 1u) % kQueue;
        if (e & m.leaf_tag) {
            // a leaf: collideTriangle(bvh[iter].bit * 3, ...) (cl:521-531)
            const uint32_t t = e & ~m.leaf_tag;
            if (t >= m.ntri) return kWalkMiss;
            float4 v0, v1, v2;
            if (!triangle_vertices(m.tri, m.vertex, m.nv, t, v0, v1, v2)) return kWalkMiss;
            if (collide_triangle(v0, v1, v2, P, V, F, p)) return kWalkHit;
        } else {
            // an internal node: collideAABB on both children, left first (cl:533-546)
            if (e >= m.ninternal) return kWalkMiss;
            const float4* w = m.wide + 4 * (size_t)e;
            const float4 a = w[0], b = w[1], c = w[2], d = w[3];
            const uint32_t el = __float_as_uint(d.x), er = __float_as_uint(d.y);
            if (collide_aabb(ray, a.x, a.y, a.z, a.w, b.x, b.y)) {
                ring.put(tail, el);
                tail = (tail + 1u) % kQueue;
                if (Ring::cap < kQueue && ((tail - head) % kQueue) > Ring::cap) return kWalkSpill;
            }
            if (collide_aabb(ray, b.z, b.w, c.x, c.y, c.z, c.w)) {
                ring.put(tail, er);
                tail = (tail + 1u) % kQueue;
                if (Ring::cap < kQueue && ((tail - head) % kQueue) > Ring::cap) return kWalkSpill;
            }
        }
    }
    return kWalkMiss;
}

__global__ void k_zero_count(uint32_t* __restrict__ spill)
{
    if (threadIdx.x == 0) spill[threadIdx.x] = 0u;
}

// One lane per particle with the LDS ring; walks that outgrow it are listed in spill[1 + k]
// (count in spill[0]) for k_collide_spill.
template <class Slot>
__global__ __launch_bounds__(kBlock) void k_collide_lds(float4* __restrict__ force, float4* __restrict__ pos,
                                                       float4* __restrict__ vel, uint32_t n, mfl_params p, Mesh m,
                                                       uint32_t* __restrict__ spill)
{
    __shared__ Slot ring[kRingLds * kBlock];
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    float4 P = pos[q], V = vel[q], F = force[q];
    LdsRing<Slot> r{ring + threadIdx.x};
    const int res = walk(r, m, p, P, V, F);
    if (res == kWalkHit) {
        pos[q] = P;
        vel[q] = V;
        force[q] = F;
    } else if (res == kWalkSpill) {
        spill[1u + atomicAdd(spill, 1u)] = q;
    }
}

// The listed particles again, each with the full 256-slot ring; a grid-stride loop over the list.
__global__ __launch_bounds__(kBlock) void k_collide_spill(float4* __restrict__ force, float4* __restrict__ pos,
                                                         float4* __restrict__ vel, mfl_params p, Mesh m,
                                                         const uint32_t* __restrict__ spill)
{
    const uint32_t count = spill[0];
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock) {
        const uint32_t q = spill[1u + k];
        float4 P = pos[q], V = vel[q], F = force[q];
        PrivateRing r;
        if (walk(r, m, p, P, V, F) == kWalkHit) {
            pos[q] = P;
            vel[q] = V;
            force[q] = F;
        }
    }
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

int BvhTrimesh::compile()
{
    if (!sort_.compile()) return 0;
    compiled_ = true;
    return 1;
}

int BvhTrimesh::compute(const float4* vertex, size_t vertex_count, const uint32_t* index, size_t index_count,
                        hipStream_t s)
{
    if (!compiled_) {
        set_error("oclBvhTrimesh: compile() first");
        return 0;
    }
    const size_t T = index_count / 3;  // cpp:56
    if (T < 2) {
        set_error("oclBvhTrimesh: at least 2 triangles are needed (index count " + std::to_string(index_count) + ")");
        return 0;
    }
    if (!vertex || !index || vertex_count == 0) {
        set_error("oclBvhTrimesh: null vertex or index buffer");
        return 0;
    }
    if (T > (1u << 30) || vertex_count > 0xFFFFFFFFull) {
        set_error("oclBvhTrimesh: mesh too large");
        return 0;
    }
    const uint32_t t = (uint32_t)T, nv = (uint32_t)vertex_count;
    count_ = 0;  // not built until this build succeeds
    root_ = 0;
    if (!part_.fit((2 * kBoxBlocks + 2) * sizeof(float4)) || !keys_.fit(T * 4) || !vals_.fit(T * 4) ||
        !nodes_.fit((2 * T - 1) * sizeof(BvhNode)) || !split_.fit(T * 4) || !range_.fit(T * sizeof(uint2)) ||
        !info_.fit(2 * sizeof(uint32_t)) || !wide_.fit((T - 1) * 4 * sizeof(float4)) || !sort_.reserve(T, 32))
        return 0;
    float4* part = part_.as<float4>();
    uint32_t* info = info_.as<uint32_t>();
    const uint32_t box_blocks = std::min<uint32_t>(kBoxBlocks, std::max<uint32_t>(1u, blocks_for(nv)));
    hipLaunchKernelGGL(k_box_partial, dim3(box_blocks), dim3(kBlock), 0, s, vertex, nv, part, info);
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(kBlock), 0, s, part, box_blocks);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(T)), dim3(kBlock), 0, s, index, vertex, nv, t, part + 2 * kBoxBlocks,
                       keys_.as<uint32_t>(), info);
    MFL_HIP(hipGetLastError());
    // cpp:108-112: stable sort of all 32 key bits; the values start as 0..T-1
    if (!sort_.compute(keys_.as<uint32_t>(), vals_.as<uint32_t>(), T, 0, 32, s, true)) return 0;
    // segment tree levels: sizes n_0 = T, n_{j+1} = ceil(n_j / 2) down to 1, entries of 2 float4
    std::vector<uint64_t> offs{0};
    std::vector<uint32_t> sizes{t};
    while (sizes.back() > 1u) {
        offs.push_back(offs.back() + sizes.back());
        sizes.push_back((sizes.back() + 1u) / 2u);
    }
    const uint32_t levels = (uint32_t)sizes.size();
    const uint64_t entries = offs.back() + sizes.back();
    if (!seg_.fit(entries * 2 * sizeof(float4)) || !seg_meta_.fit(levels * (sizeof(uint64_t) + sizeof(uint32_t))))
        return 0;
    uint64_t* d_offs = seg_meta_.as<uint64_t>();
    uint32_t* d_sizes = reinterpret_cast<uint32_t*>(d_offs + levels);
    seg_host_.resize(levels * 3);
    std::memcpy(seg_host_.data(), offs.data(), levels * sizeof(uint64_t));
    std::memcpy(reinterpret_cast<char*>(seg_host_.data()) + levels * sizeof(uint64_t), sizes.data(),
                levels * sizeof(uint32_t));
    MFL_HIP(hipMemcpyAsync(d_offs, seg_host_.data(), levels * (sizeof(uint64_t) + sizeof(uint32_t)),
                           hipMemcpyHostToDevice, s));
    float4* seg = seg_.as<float4>();
    hipLaunchKernelGGL(k_nodes, dim3(blocks_for(T)), dim3(kBlock), 0, s, keys_.as<uint32_t>(), vals_.as<uint32_t>(),
                       index, vertex, nv, t, nodes(), split_.as<uint32_t>(), seg);
    hipLaunchKernelGGL(k_link, dim3(blocks_for(2 * T - 1)), dim3(kBlock), 0, s, keys_.as<uint32_t>(),
                       split_.as<uint32_t>(), t, nodes(), range_.as<uint2>(), info);
    for (uint32_t k = 0; k + 1 < levels; k += 8) {
        const uint32_t up = std::min<uint32_t>(8u, levels - 1u - k);
        hipLaunchKernelGGL(k_seg_up, dim3(blocks_for(sizes[k])), dim3(kBlock), 0, s, seg, d_offs, d_sizes, k, up);
    }
    hipLaunchKernelGGL(k_node_boxes, dim3(blocks_for(T)), dim3(kBlock), 0, s, seg, d_offs, range_.as<uint2>(), t,
                       nodes());
    hipLaunchKernelGGL(k_wide, dim3(blocks_for(T)), dim3(kBlock), 0, s, nodes(), t, 2 * T - 1 <= 0xFFFFu ? 0x8000u : 0x80000000u,
                       wide_.as<float4>());
    MFL_HIP(hipGetLastError());
    uint32_t host[2] = {0, 0};
    MFL_HIP(hipMemcpyAsync(host, info, sizeof(host), hipMemcpyDeviceToHost, s));
    MFL_HIP(hipStreamSynchronize(s));
    if (host[1] == 1u) {
        set_error("oclBvhTrimesh: a triangle index is not below the vertex count");
        return 0;
    }
    if (host[1] != 0u) {
        set_error("oclBvhTrimesh: the tree links are incomplete");
        return 0;
    }
    if (host[0] >= 2 * T - 1) {
        set_error("oclBvhTrimesh: tree root out of range");
        return 0;
    }
    count_ = (uint32_t)(2 * T - 1);
    root_ = host[0];
    return 1;
}

int BvhTrimesh::collide(float4* force, float4* pos, float4* vel, size_t n, const mfl_params& p,
                        const uint32_t* triangle, size_t index_count, const float4* vertex, size_t vertex_count,
                        hipStream_t s)
{
    if (n == 0) return 1;
    if (count_ == 0) {
        set_error("clCollideBVH: the BVH has not been built");
        return 0;
    }
    if (!force || !pos || !vel || !triangle || !vertex) {
        set_error("clCollideBVH: null buffer");
        return 0;
    }
    if (n > 0xFFFFFFFEull || vertex_count > 0xFFFFFFFFull || index_count / 3 > 0xFFFFFFFFull) {
        set_error("clCollideBVH: sizes beyond 32-bit indices");
        return 0;
    }
    if (!spill_.fit((n + 1) * sizeof(uint32_t))) return 0;
    const Mesh m{wide_.as<float4>(), (count_ + 1u) / 2u - 1u, root_, leaf_tag(), triangle, (uint32_t)(index_count / 3),
                 vertex, (uint32_t)vertex_count};
    uint32_t* spill = spill_.as<uint32_t>();
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, spill);
    if (count_ <= 0xFFFFu)
        hipLaunchKernelGGL(k_collide_lds<uint16_t>, dim3(blocks_for(n)), dim3(kBlock), 0, s, force, pos, vel,
                           (uint32_t)n, p, m, spill);
    else
        hipLaunchKernelGGL(k_collide_lds<uint32_t>, dim3(blocks_for(n)), dim3(kBlock), 0, s, force, pos, vel,
                           (uint32_t)n, p, m, spill);
    hipLaunchKernelGGL(k_collide_spill, dim3(std::min<uint32_t>(blocks_for(n), 1024u)), dim3(kBlock), 0, s, force,
                       pos, vel, p, m, spill);
    MFL_HIP(hipGetLastError());
    return 1;
}

}  // namespace mfl
