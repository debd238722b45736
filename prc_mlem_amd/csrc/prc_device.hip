// prc_device.hip -- gfx950 kernels of the positron-range-correction path and the extern "C"
// layer declared in include/prc.h.
//
// The system matrix A of the reference (prc.cu:167-267) is fixed for a reconstruction: it
// depends on density, geometry and isotope only. Instead of re-marching every ray in each
// forward and adjoint pass (prc.cu:178-230), the normalisation pass marches every ray once
// and stores each per-step weight in a "row-plane" layout:
//
//   for each X row (y, z) holding active voxels, x in [xa, xa + len):
//     for each ray p (p = 0: the 125-tap near block, p = 1 + r: far ray r):
//       cnt(row, p) steps k = kfirst(p) .. kfirst(p) + cnt - 1, each a plane of len weights [x].
//
// For a fixed ray and step, the DDA voxel offset (dx, dy, dz) is the same for every source
// (prc.cu:118-129, 140-142 depend on the voxel spacing only), so
//   adjoint: out_s = sum_{p,k} w[row(s)][p][k][x_s] * in[s + D(p,k)] / W_s   (stream the planes)
//   forward: out_t = sum_{p,k} w[row(t - D)][p][k][x_t - dx] * f/W[t - D]    (gather the planes)
// are both gathers with a wave-uniform base address plus the lane index: no atomics, a fixed
// summation order, and A^T is exactly the transpose of A (same stored weights).
//
// All kernels take a table of volume descriptors and a list of work items (vol, row, x0, n)
// so one launch serves any number of volumes (the batched path).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "prc.h"

namespace {

constexpr int kBins = PRC_RANGE_BINS;    // prc.cu:61
constexpr int kNearRays = PRC_NEAR_RAYS; // prc.cu:104
static_assert(kNearRays % 2 == 0, "the near march takes rays in pairs");
constexpr int kNQ = 125;                 // 5x5x5 near block (prc.cu:103)
constexpr int kNearMax = 8;              // a near ray visits at most 7 block voxels
constexpr float kMinDensity = 1.0e-3f;   // prc.cu:102
// Density outside the grid (the padded volume's halo and margins): a quiet NaN. No comparison
// holds for it, so such a voxel is never tissue and never counts as a visit, and
// fminf(fmaf(NaN, seg, w), cutoff) = cutoff (fminf returns its non-NaN operand), so a step onto
// it leaves the lane dead without a deposit: the grid exit of prc.cu:163.
constexpr float kHalo = __builtin_nanf("");
constexpr int kWave = 64;
constexpr int kItemsPerBlock = 4;        // 256-thread blocks, one 64-voxel row segment per wave
constexpr int kFarGroup = 8;             // far-march steps whose density loads are issued together
constexpr int kCountGroup = 16;          // the same for the count pass
constexpr int kMarchGroup = kCountGroup > kFarGroup ? kCountGroup : kFarGroup;

// ---------------------------------------------------------------------------------------
// Isotope model (prc.cu:38-57): Paneque-Yunta et al., Phys. Med. Biol. 70 (2025) 165013.
struct Branch {
    float w, a, b, c, f;
};
struct Isotope {
    const char* name;
    int n;
    Branch br[3];
};
const Isotope kIsotopes[8] = {
    {"11C", 1, {{1.000f, 0.777f, 0.468f, 0.217f, 0.010f}}},
    {"13N", 1, {{1.000f, 0.537f, 0.347f, 0.193f, 0.050f}}},
    {"15O", 1, {{1.000f, 0.295f, 0.221f, 0.133f, 0.118f}}},
    {"18F", 1, {{1.000f, 1.612f, 0.703f, 0.526f, 0.011f}}},
    {"64Cu", 1, {{1.000f, 1.494f, 0.000f, 0.631f, 0.251f}}},
    {"68Ga", 2, {{0.014f, 1.048f, 0.000f, 0.475f, 0.271f}, {0.986f, 0.289f, 0.137f, 0.159f, 0.322f}}},
    {"82Rb", 2, {{0.136f, 0.191f, 0.084f, 0.114f, 0.408f}, {0.857f, 0.135f, 0.063f, 0.087f, 0.451f}}},
    {"124I", 3, {{0.013f, 1.007f, 0.000f, 0.421f, 0.443f}, {0.515f, 0.412f, 0.000f, 0.210f, 0.458f},
                 {0.471f, 0.241f, 0.000f, 0.150f, 0.369f}}},
};

// The smallest branch absorbs the FP32 rounding of the published weights (prc.cu:577-584).
Isotope renormalised(int idx) {
    Isotope m = kIsotopes[idx];
    float wsum = 0.0f;
    int smallest = 0;
    for (int b = 0; b < m.n; ++b) {
        wsum += m.br[b].w;
        if (m.br[b].w < m.br[smallest].w) smallest = b;
    }
    m.br[smallest].w += 1.0f - wsum;
    return m;
}

// S(r) = sum_b w_b exp(-[(a r)^(1+f) + (b r)^2.5 + (c r)^3.5]) (prc.cu:65-72), evaluated
// without contraction, as the reference's host table build does.
__host__ __device__ inline float survival(const Branch* br, int n, float r) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {  // at most three branches; a fixed trip count keeps br in registers
        if (i >= n) break;
        const float x = br[i].b * r, y = br[i].c * r;
        const float p = powf(br[i].a * r, 1.0f + br[i].f);
        const float q = x * x * sqrtf(x);
        const float u = y * y * y * sqrtf(y);
        s = s + br[i].w * expf(-((p + q) + u));
    }
    return s;
}

// prc.cu:83-91 (host, FP32).
float range_cutoff_host(const Isotope& m, float fraction) {
    float lo = 0.0f, hi = 1.0f;
    while (1.0f - survival(m.br, m.n, hi) < fraction && hi < 256.0f) hi *= 2.0f;
    for (int it = 0; it < 64; ++it) {
        const float mid = 0.5f * (lo + hi);
        if (1.0f - survival(m.br, m.n, mid) < fraction) lo = mid;
        else hi = mid;
    }
    return hi;
}

// ---------------------------------------------------------------------------------------
// Device data layout.
// A row's active x range [xa, xa + len) is cut into segments of up to 64 voxels starting at xa:
// segment s covers [xa + 64 s, xa + 64 s + min(64, len - 64 s)). One wave handles one segment,
// and every (segment, ray) owns one contiguous run of planes, so a wave streams memory no other
// wave touches. Unit index of segment s of row r: r * nseg + s (nseg = ceil(nx / 64)).
struct RowInfo {        // one per (segment, ray); 32 bytes, one scalar load
    long long off;      // element offset of this segment's planes for this ray in the weight buffer
    int cnt;            // stored steps (planes)
    int pad2;
    int xa, len;        // the segment's x range [xa, xa + len) = plane width (len 0: empty segment)
    int pad0, pad1;
};
// The planes of one segment in one copy of the matrix: the runs of all its rays, in ray order
// (near block first), follow each other with no gap, so they form one sequence of planes of the
// segment's width; the shift list holds the gather shift of each of those planes.
struct PlaneRun {       // 16 bytes, one scalar load
    long long off;      // first entry of the segment's planes in the weight buffer
    int planes;         // planes of all its rays
    int dbase;          // first entry of its shifts in the shift list
};
// Where the target-major planes of one segment for one ray lie (a compact, ray-major copy of the
// target layout's offsets and counts: neighbouring rows of one ray are neighbouring entries).
struct TargetRun {      // 16 bytes
    long long off;      // first entry of the run
    int cnt;            // planes
    int pad;
};
struct FarStep {        // march table of a far ray (16 bytes, read with wave-uniform scalar loads)
    float seg;          // path length [mm] inside the step's voxel (t_next - t, prc.cu:150)
    unsigned soff;      // 2^31 + 4 * (padded linear offset of the step's voxel from the source): the
                        // scalar offset of the step's density load (RhoReader::step)
    short dx, dy, dz;   // the step's voxel relative to the source (the deposit's target)
    short pad;
};
struct NearStep {
    float seg;
    int dp;
    int q;              // near-block bin 0..124
    int loc;            // index of the bin inside its Manhattan shell (step s visits shell s)
};

// Near-block bins grouped by Manhattan distance from the source voxel. A DDA step moves one
// voxel along one axis away from the source, so step s of a near ray always lands in shell s
// (sizes 1, 6, 18, 32, 36, 24, 8): per-step accumulators stay small enough for register
// indexing.
struct ShellMap {
    int shell[kNQ];
    int local[kNQ];
};
constexpr int iabs_c(int v) { return v < 0 ? -v : v; }
constexpr ShellMap make_shell_map() {
    ShellMap m{};
    int count[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < kNQ; ++q) {
        const int sh = iabs_c(q % 5 - 2) + iabs_c(q / 5 % 5 - 2) + iabs_c(q / 25 - 2);
        m.shell[q] = sh;
        m.local[q] = count[sh]++;
    }
    return m;
}
constexpr ShellMap kShellMap = make_shell_map();

struct Vol {            // device descriptor of one volume
    int nx, ny, nz, n_rays;
    float sx, sy, sz, cutoff, support_rho, air_cap, scale;
    int nb, exact, L, LP;
    int LF;             // step-table stride: L + kFarGroup rounded up to kMarchGroup, zero steps past L
    Branch br[3];
    size_t vbase;       // offset of this volume in concatenated voxel vectors
    const float* rho_p; // (nx+2)(ny+2)(nz+2), clamped density, kHalo (NaN) halo
    const float* rho_base;  // start of the allocation holding rho_p and its kHalo margins
    long long rho_bytes;    // size of that allocation
    const float* table; // kBins + 1
    const unsigned char* support;
    float* W;           // per voxel
    const int* row_xa;  // [ny*nz]
    const int* row_len; // [ny*nz], 0 = no active voxel
    int nseg;           // segment slots per row, ceil(nx / 64)
    int box[6];         // bounding box of the support: x0, x1, y0, y1, z0, z1 (inclusive; empty: x0 > x1)
    RowInfo* rinfo;     // [ny*nz*nseg][n_rays + 1] source-major layout (planes over source x)
    RowInfo* tinfo;     // [ny*nz*nseg][n_rays + 1] target-major layout (planes over target x)
    const TargetRun* tmaj;  // [n_rays + 1][ny*nz*nseg] the target-major runs, ray-major
    void* wbuf;         // source-major weights (float or __half): the adjoint's stream
    void* tbuf;         // target-major weights: the forward's stream
    const PlaneRun* sruns;  // [ny*nz*nseg] source-major plane sequence of each segment
    const PlaneRun* truns;  // [ny*nz*nseg] target-major plane sequence of each segment
    const int* slist;       // source-major shifts +D (the adjoint gathers at s + D)
    const int* tlist;       // target-major shifts -D (the forward gathers at t - D)
    const FarStep* fstep;   // [n_rays][LF]
    const unsigned* fsoa;   // [n_rays][2][LF]: the steps' seg bits, then their soff (a group's steps
                            // in two wide scalar loads)
    const int4* pstep;      // [n_rays + 1][LP] (dx, dy, dz) of each step; p = 0 is the near block
    const int* plin;        // [n_rays + 1][LP] linear voxel offset (dz*ny + dy)*nx + dx of each step
    const int* kfirst;      // [n_rays + 1] first stored step
    const NearStep* nstep;  // [kNearRays][kNearMax]
    const int* nlen;        // [kNearRays]
};

// Solver scalars of one volume (device).
struct Scal {
    double support_sum;
    double sum;          // last FP64 sum
    double hh, hy;       // Biggs-Andrews dot products
    float eps, lambda, scale, beta, gamma;
    int accel;
    // Mixed form: powers of two applied to the FP16 copies that the projector passes gather, set
    // every iteration from the vector's largest value so that it lands in [2^13, 2^14): the copy
    // can neither overflow (65504) nor drop values above 2^-28 of its maximum into subnormals.
    // hscale scales f / W (the forward's output is scaled back exactly in k_ratio), rscale the
    // ratio (the adjoint's output is scaled back exactly by rinv = 1 / rscale in k_update).
    float hscale, rscale, rinv;
};

// Loads and stores through the global address space. Pointers kept in descriptors are generic
// pointers to the compiler; casting them lets wave-uniform reads become scalar loads and lets
// per-lane accesses use global (not flat) instructions.
#define GLOBAL_AS __attribute__((address_space(1)))
#define CONST_AS __attribute__((address_space(4)))  // global memory that no kernel writes: scalar loads
typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));

template <class T>
__device__ inline T gload(const T* p) {
    if constexpr (sizeof(T) == 32) return __builtin_bit_cast(T, *(const GLOBAL_AS v8i*)p);
    else if constexpr (sizeof(T) == 16) return __builtin_bit_cast(T, *(const GLOBAL_AS v4i*)p);
    else if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, *(const GLOBAL_AS v2i*)p);
    else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, *(const GLOBAL_AS int*)p);
    else if constexpr (sizeof(T) == 2) return __builtin_bit_cast(T, *(const GLOBAL_AS unsigned short*)p);
    else return __builtin_bit_cast(T, *(const GLOBAL_AS unsigned char*)p);
}

template <class T>
__device__ inline void gstore(T* p, T v) {
    if constexpr (sizeof(T) == 4) *(GLOBAL_AS int*)p = __builtin_bit_cast(int, v);
    else if constexpr (sizeof(T) == 2) *(GLOBAL_AS unsigned short*)p = __builtin_bit_cast(unsigned short, v);
    else *(GLOBAL_AS T*)p = v;
}

template <class T> struct WeightIO;
template <> struct WeightIO<float> {
    __device__ static float load(const float* p) { return gload(p); }
    __device__ static float store_value(float v) { return v; }
    __device__ static float to_float(float v) { return v; }
};
template <> struct WeightIO<__half> {
    __device__ static float load(const __half* p) { return __half2float(gload(p)); }
    __device__ static __half store_value(float v) { return __float2half_rn(v); }
    __device__ static float to_float(__half v) { return __half2float(v); }
};

// Buffer store of one stored weight (the bits as they are) at byte offset voff + soff.
template <class WT>
__device__ inline void buffer_store_weight(WT v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
    if constexpr (sizeof(WT) == 2)
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, voff, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// Buffer descriptor over `bytes` bytes from a wave-uniform base: an access then costs a 32-bit
// offset instead of 64-bit address arithmetic, and accesses at offsets past `bytes` are dropped
// (stores) or read 0 (loads) by the range check. Descriptor word 3 as for gfx9-family chips.
__device__ inline __amdgpu_buffer_rsrc_t buffer_of(const void* base, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)min(bytes, 0x7fffffffLL),
                                             0x00020000);
}
// Bias of descriptors whose base is moved back by 2^31 bytes from the data they address: an offset
// of either sign within +-2^31 bytes of the data becomes a nonnegative 32-bit buffer offset. Such
// descriptors span the whole 32-bit offset range (no range clamp).
constexpr unsigned kGatherBias = 0x80000000u;

// Density reads of the marches: buffer loads whose address is split into the lane's part (the
// vector offset, fixed per lane) and the step's wave-uniform part plus kGatherBias (the scalar
// offset, straight from a scalar step-table load), over a descriptor moved back by kGatherBias
// bytes. Both offsets are nonnegative and their sum, the byte offset from the moved base, stays
// below 2^32 (the hardware adds them without wrapping; a volume's padded density fits in 2^31
// bytes), so a step's address costs no vector instruction.
struct RhoReader {
    __amdgpu_buffer_rsrc_t r;
    int lane_off;
    // index: the lane's source in the padded density, counted from rho_p.
    __device__ RhoReader(const float* rho_p, size_t index)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)((const char*)rho_p - kGatherBias), (short)0, (int)0xffffffffu,
                                              0x00020000)),
          lane_off((int)(4u * (unsigned)index)) {}
    // Density at padded offset dp (wave-uniform) from the lane's source.
    __device__ float operator()(int dp) const { return step(kGatherBias + 4u * (unsigned)dp); }
    // The same with the scalar offset kGatherBias + 4 * dp given (a far step's soff).
    __device__ float step(unsigned soff) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane_off, (int)soff, 0));
    }
};

// Table lookup of prc.cu:74-80 (u = w * scale; i = min(int(u), kBins - 1); lo + min(u - i, 1) *
// (hi - lo)) on the LDS table with its extra entry tab[kBins + 1] = tab[kBins], without the clamps.
// The march only asks for 0 <= w <= cutoff, so u <= kBins to within one rounding: for u < kBins
// both forms take the same bin and fraction; for u >= kBins the reference form returns
// lo + 1 * (T[kBins] - lo) with lo = T[kBins - 1], which is T[kBins] exactly (neighbouring
// survivals are within a factor of two, so the difference is exact), and this one T[kBins] + f * 0.
__device__ inline float march_lookup(const float* tab, float scale, float w) {
    const float u = w * scale;
    const int i = (int)u;
    const float lo = tab[i], hi = tab[i + 1];
    return fmaf(__builtin_amdgcn_fractf(u), hi - lo, lo);  // fract(u) = u - i exactly for 0 <= u < 2^23
}

// Index of the calling wave inside its block, known to the compiler as wave-uniform so that
// item and descriptor loads become scalar loads.
__device__ inline int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave)); }

__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ inline long long wave_sum_ll(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ inline double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// ---------------------------------------------------------------------------------------
// Tables built on the device.

__global__ void k_range_table(float* table, Branch b0, Branch b1, Branch b2, int nb, float scale) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > kBins) return;
    const Branch br[3] = {b0, b1, b2};
    table[j] = survival(br, nb, (float)j / scale);  // prc.cu:625-633
}

// The ray set of prc.cu:118-129: mirrored pairs, Fibonacci hemisphere directions and R3 offsets.
// sincospif is evaluated as the correctly rounded sin/cos of pi*arg via FP64.
__device__ void ray_dir(int k, int n, float3* start, float3* dir) {
    const unsigned m = (unsigned)k >> 1;
    const float sign = (k & 1) ? -1.0f : 1.0f;
    const float z = 1.0f - ((float)m + 0.5f) / (float)((n + 1) >> 1);
    const float r = sqrtf(fmaxf(0.0f, fmaf(-z, z, 1.0f)));
    const float to_unit = 2.3283064e-10f;  // 2^-32
    const float arg = 2.0f * to_unit * (float)(m * 2654435769u);
    const double pa = 3.14159265358979323846 * (double)arg;
    const float sn = (float)sin(pa), cs = (float)cos(pa);
    *dir = make_float3(sign * r * cs, sign * r * sn, sign * z);
    *start = make_float3(sign * fmaf(to_unit, (float)(0x80000000u + m * 3518319155u), -0.5f),
                         sign * fmaf(to_unit, (float)(0x80000000u + m * 2882110345u), -0.5f),
                         sign * fmaf(to_unit, (float)(0x80000000u + m * 2360945575u), -0.5f));
}

// The ray set itself, for validation: start offsets then directions of rays 0 .. n - 1.
__global__ void k_ray_set(int n, float* __restrict__ out) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    float3 x, d;
    ray_dir(id, n, &x, &d);
    float* o = out + (size_t)6 * id;
    o[0] = x.x, o[1] = x.y, o[2] = x.z, o[3] = d.x, o[4] = d.y, o[5] = d.z;
}

// Geometric voxel walk of one ray (prc.cu:137-163 without the density): the voxel offset and
// the in-voxel path length of every step are the same for all sources.
struct Walk {
    int di, dj, dk, i, j, k;
    float tx, ty, tz, dtx, dty, dtz, t;
    __device__ Walk(float3 x, float3 d, float sx, float sy, float sz) {
        const float big = 1.0e30f;
        di = d.x > 0.0f ? 1 : -1, dj = d.y > 0.0f ? 1 : -1, dk = d.z > 0.0f ? 1 : -1;
        tx = d.x != 0.0f ? fmaf(0.5f, (float)di, -x.x) * sx / d.x : big;
        ty = d.y != 0.0f ? fmaf(0.5f, (float)dj, -x.y) * sy / d.y : big;
        tz = d.z != 0.0f ? fmaf(0.5f, (float)dk, -x.z) * sz / d.z : big;
        dtx = d.x != 0.0f ? sx / fabsf(d.x) : big;
        dty = d.y != 0.0f ? sy / fabsf(d.y) : big;
        dtz = d.z != 0.0f ? sz / fabsf(d.z) : big;
        i = j = k = 0;
        t = 0.0f;
    }
    __device__ bool in_block() const { return abs(i) <= 2 && abs(j) <= 2 && abs(k) <= 2; }
    __device__ float step() {  // returns the segment of the current voxel and advances
        const float tn = fminf(tx, fminf(ty, tz));
        const float seg = tn - t;
        t = tn;
        if (tn == tx) i += di, tx += dtx;
        else if (tn == ty) j += dj, ty += dty;
        else k += dk, tz += dtz;
        return seg;
    }
};

__device__ inline int linear_offset(const Vol& V, int i, int j, int k) {
    const long long v = ((long long)k * V.ny + j) * V.nx + i;  // steps that are ever used stay inside the grid
    return (int)max(min(v, (long long)INT_MAX), (long long)INT_MIN);
}

__global__ void k_rays(Vol V, FarStep* fstep, unsigned* fsoa, int4* pstep, int* plin, int* kfirst, NearStep* nstep,
                       int* nlen) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxp = V.nx + 2, nyp = V.ny + 2;
    if (id < V.n_rays) {
        float3 x, d;
        ray_dir(id, V.n_rays, &x, &d);
        Walk w(x, d, V.sx, V.sy, V.sz);
        int kf = -1;
        for (int s = 0; s < V.L; ++s) {
            if (kf < 0 && !w.in_block()) kf = s;
            const int i = w.i, j = w.j, k = w.k;
            pstep[(size_t)(1 + id) * V.LP + s] = make_int4(i, j, k, 0);
            plin[(size_t)(1 + id) * V.LP + s] = linear_offset(V, i, j, k);
            const float seg = w.step();
            const unsigned soff = kGatherBias + 4u * (unsigned)((k * nyp + j) * nxp + i);
            fstep[(size_t)id * V.LF + s] = FarStep{seg, soff, (short)i, (short)j, (short)k, 0};
            fsoa[(size_t)id * 2 * V.LF + s] = __float_as_uint(seg);
            fsoa[(size_t)id * 2 * V.LF + V.LF + s] = soff;
        }
        // Steps past L: no move and no length (the source voxel itself).
        for (int s = V.L; s < V.LF; ++s) {
            fstep[(size_t)id * V.LF + s] = FarStep{0.0f, kGatherBias, 0, 0, 0, 0};
            fsoa[(size_t)id * 2 * V.LF + s] = 0u;
            fsoa[(size_t)id * 2 * V.LF + V.LF + s] = kGatherBias;
        }
        kfirst[1 + id] = kf < 0 ? V.L : kf;
    } else if (id < V.n_rays + kNearRays) {
        const int r = id - V.n_rays;
        float3 x, d;
        ray_dir(r, kNearRays, &x, &d);
        Walk w(x, d, V.sx, V.sy, V.sz);
        int s = 0;
        for (; s < kNearMax && w.in_block(); ++s) {
            const int i = w.i, j = w.j, k = w.k;
            const float seg = w.step();
            const int q = ((k + 2) * 5 + j + 2) * 5 + i + 2, sh = abs(i) + abs(j) + abs(k);
            int loc = 0;
            for (int q2 = 0; q2 < q; ++q2) loc += abs(q2 % 5 - 2) + abs(q2 / 5 % 5 - 2) + abs(q2 / 25 - 2) == sh;
            nstep[r * kNearMax + s] = NearStep{seg, (k * nyp + j) * nxp + i, q, loc};
        }
        nlen[r] = s;
    } else if (id < V.n_rays + kNearRays + kNQ) {
        const int q = id - V.n_rays - kNearRays;
        pstep[q] = make_int4(q % 5 - 2, q / 5 % 5 - 2, q / 25 - 2, 0);
        plin[q] = linear_offset(V, q % 5 - 2, q / 5 % 5 - 2, q / 25 - 2);
        if (q == 0) kfirst[0] = 0;
    }
}

// ---------------------------------------------------------------------------------------
// Normalisation march (prc.cu:178-237 with NormOp): one wave = 64 consecutive sources of a row.
// Per step the DDA offset is wave-uniform, so density reads are (uniform base + lane).

// March state of one lane. A dead lane holds w = cutoff and surv = S(cutoff): every further step
// then leaves it unchanged and deposits nothing, so no separate alive flag is carried.
struct Lane {
    float w, surv, air;
};

// March parameters of a volume, copied out of the descriptor once per wave: the kernels store
// through pointers the compiler cannot tell apart from the descriptor, which would otherwise
// reload these in every step.
struct MarchConst {
    float cutoff, scale, support_rho, air_cap;
    int nb;
    Branch br[3];
    float s_cut;  // S(cutoff) as the step evaluates it: the survival a dead lane keeps
    template <bool EXACT>
    __device__ MarchConst(const Vol& V, const float* tab, std::integral_constant<bool, EXACT>)
        : cutoff(V.cutoff), scale(V.scale), support_rho(V.support_rho), air_cap(V.air_cap), nb(V.nb),
          br{V.br[0], V.br[1], V.br[2]} {
        s_cut = EXACT ? survival(br, nb, cutoff) : march_lookup(tab, scale, cutoff);
    }
    __device__ Lane start(bool act) const { return Lane{act ? 0.0f : cutoff, act ? 1.0f : s_cut, 0.0f}; }
    __device__ bool alive(const Lane& st) const { return st.w < cutoff; }
};

// One DDA step of prc.cu:146-163 for the current lane; returns the deposit (0 if none). Written
// with selects instead of branches: the ray leaves the grid at the kHalo halo, dies in air after
// air_cap mm, and dies when the range reaches the cutoff. A dead lane (w = cutoff) computes
// wn = cutoff and sn = S(cutoff) = surv, so it deposits nothing and stays dead.
template <bool EXACT>
__device__ inline float march_step(Lane& st, float rho, float seg, bool deposit_here, const MarchConst& V,
                                   const float* tab, unsigned& visits) {
    visits += (rho < INFINITY && st.w < V.cutoff) ? 1u : 0u;  // a voxel of the grid (not kHalo)
    const float wn = fminf(fmaf(rho, seg, st.w), V.cutoff);
    const float sn = EXACT ? survival(V.br, V.nb, wn) : march_lookup(tab, V.scale, wn);
    const bool tissue = rho >= V.support_rho;
    const float air = tissue ? 0.0f : st.air + seg;
    // surv - sn > 0 exactly when surv > sn, so the max is the reference's guarded difference.
    // A kHalo voxel is not tissue, so it deposits nothing.
    const float dep = (tissue && deposit_here) ? fmaxf(st.surv - sn, 0.0f) : 0.0f;
    // In tissue air = 0 and air_cap > 0, so the air test needs no tissue test. A step onto the
    // halo needs no test either: wn = fminf(fmaf(NaN, seg, w), cutoff) = cutoff, so
    // sn = S(cutoff) = s_cut, the state of a dead lane.
    const bool dies = air > V.air_cap;
    st.w = dies ? V.cutoff : wn;
    st.surv = dies ? V.s_cut : sn;
    st.air = air;
    return dep;
}

// The march's copy of the table in LDS carries one extra entry, tab[kBins + 1] = tab[kBins], so
// that march_lookup needs no clamps.
constexpr int kTabLds = kBins + 2;
__device__ inline void load_table(float* tab, const Vol& V) {
    for (int j = threadIdx.x; j < kTabLds; j += blockDim.x) tab[j] = gload(V.table + min(j, kBins));
    __syncthreads();
}

// Target-major slots of a wave's deposits. The wave's lanes hit targets tx = tx0 + lane of row t,
// which fall into at most two segments of that row; lanes below `thr` use the first. For plane k
// the slot of a lane is tbuf[b + lane] with b from its segment; `ok` bit 0 / bit 1 says whether
// the segment holds plane k (it always does where a deposit lands; the check keeps every store
// inside its own run regardless). unit0 points at segment 0 of row t for the ray.
struct TargetPair {
    long long b0, b1;
    int thr, ok;
};
__device__ inline TargetPair target_pair(const RowInfo* unit0, int R1, int xa_t, int len_t, int tx0, int k) {
    TargetPair tp{0, 0, kWave, 0};
    const int rel = tx0 - xa_t;
    const int s0 = rel >= 0 ? rel / kWave : -((kWave - 1 - rel) / kWave);  // floor
    const int ns = (len_t + kWave - 1) / kWave;
    tp.thr = kWave * (s0 + 1) - rel;  // 1 .. 64
    if (s0 >= 0 && s0 < ns) {
        const RowInfo e = gload(unit0 + (size_t)s0 * R1);
        tp.ok |= k < e.cnt ? 1 : 0;
        tp.b0 = e.off - e.xa + (long long)k * e.len + tx0;
    }
    if (s0 + 1 >= 0 && s0 + 1 < ns) {
        const RowInfo e = gload(unit0 + (size_t)(s0 + 1) * R1);
        tp.ok |= k < e.cnt ? 2 : 0;
        tp.b1 = e.off - e.xa + (long long)k * e.len + tx0;
    }
    return tp;
}
// The same from the ray-major runs: row0 points at segment 0 of row t for the ray; the segments'
// x ranges follow from the row's active range.
__device__ inline TargetPair target_pair_run(const TargetRun* row0, int xa_t, int len_t, int tx0, int k) {
    TargetPair tp{0, 0, kWave, 0};
    const int rel = tx0 - xa_t;
    const int s0 = rel >= 0 ? rel / kWave : -((kWave - 1 - rel) / kWave);  // floor
    const int ns = (len_t + kWave - 1) / kWave;
    tp.thr = kWave * (s0 + 1) - rel;  // 1 .. 64
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int sg = s0 + h;
        if (sg >= 0 && sg < ns) {
            const TargetRun e = gload(row0 + sg);
            const int exa = xa_t + kWave * sg, elen = min(kWave, len_t - kWave * sg);
            tp.ok |= k < e.cnt ? 1 << h : 0;
            (h ? tp.b1 : tp.b0) = e.off - exa + (long long)k * elen + tx0;
        }
    }
    return tp;
}
__device__ inline bool target_slot(const TargetPair& tp, int lane, long long& idx) {
    const bool hi = lane >= tp.thr;
    idx = (hi ? tp.b1 : tp.b0) + lane;
    return ((tp.ok >> (hi ? 1 : 0)) & 1) != 0;
}

// The target-major slots of one step of a far ray as the fill pass uses them, staged in LDS by the
// lane that resolved them and read back by every lane (24 bytes). The slot of lane l is b0 + l,
// plus delta for l >= thr (the second segment); lanes below lo_lim or from hi_start on hold the
// step's plane. b0 and delta are chosen so that a missing segment needs no other case.
struct FillStep {
    long long b0;
    int delta, thr, lo_lim, hi_start;
};
__device__ inline FillStep fill_step(const TargetPair& tp) {
    const long long b0 = (tp.ok & 1) ? tp.b0 : tp.b1, b1 = (tp.ok & 2) ? tp.b1 : b0;
    return FillStep{b0, (int)(b1 - b0), tp.thr, (tp.ok & 1) ? tp.thr : 0, (tp.ok & 2) ? tp.thr : kWave};
}

// The 2048 near rays of one source per lane (prc.cu:193-204), accumulated into registers: four
// arrays of exactly 32 bins, the size of a register tuple, so that none is padded to the next one.
// Each is small enough to stay in registers under indexing and large enough to be indexed rather
// than selected element by element; the bin index of a step is wave-uniform. Layout:
//   a[0] shell 0, a[1..6] shell 1, a[7..24] shell 2, a[25..28] bins 32-35 of shell 4;
//   b shell 3; c bins 0-31 of shell 4; d[0..23] shell 5, d[24..31] shell 6.
// rho_at(dp) returns the density at padded offset dp from the lane's source. COUNT adds the voxel
// visits (a statistic) to `visits`.
#define NEAR_ACCUMULATORS float na[32], nb[32], nc[32], nd[32]
#define NEAR_ARGS na, nb, nc, nd
template <bool EXACT, bool COUNT, class RhoAt>
__device__ __attribute__((always_inline)) inline void near_rays(const Vol& V, const float* tab, bool act, RhoAt&& rho_at,
                                                              float (&a)[32], float (&b)[32], float (&c)[32],
                                                              float (&d)[32], unsigned& visits) {
#pragma unroll
    for (int i = 0; i < 32; ++i) a[i] = b[i] = c[i] = d[i] = 0.0f;
    const MarchConst M(V, tab, std::integral_constant<bool, EXACT>());
    // Two rays at a time: their steps are independent, so the dependent chain of one (depth,
    // table lookup, survival) overlaps the other's. Each bin still receives the rays' deposits in
    // ray order (a ray visits a voxel once), so the sums are those of one ray after another.
    // Steps past a ray's length have a zero step entry (seg 0, dp 0, bin 0): w stays, so sn = surv
    // and they deposit exactly 0 into bin 0 of their shell, whatever density they read; an
    // inactive lane is dead from the start. Only when counting visits do such steps read kHalo
    // (which counts no visit).
    for (int r = 0; r < kNearRays; r += 2) {
        int len[2];
        NearStep ns[2][7];
        float rv[2][7];
        // The pair's step records (zero past a ray's length): 16-byte scalar loads at immediate
        // offsets from one base per pair.
        const NearStep* np = V.nstep + r * kNearMax;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            len[h] = gload(V.nlen + r + h);
#pragma unroll
            for (int s = 0; s < 7; ++s) {
                const v4i e = *(const CONST_AS v4i*)(np + h * kNearMax + s);
                ns[h][s] = NearStep{__int_as_float(e[0]), e[1], e[2], e[3]};
            }
        }
        // The whole ray stays inside the 5x5x5 block (the kHalo margin around the padded density
        // keeps every address in bounds). Densities of steps 0-3 are loaded up front, those of
        // steps 4-6 only for ray pairs that reach them, while step 3 runs.
        auto load_rho = [&](int s0, int s1) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int s = s0; s < s1; ++s)
                    rv[h][s] = !COUNT ? rho_at(ns[h][s].dp) : (s < len[h] && act) ? rho_at(ns[h][s].dp) : kHalo;
        };
        load_rho(0, 4);
        Lane st[2] = {M.start(act), M.start(act)};
        auto step = [&](int h, int s) { return march_step<EXACT>(st[h], rv[h][s], ns[h][s].seg, true, M, tab, visits); };
#pragma unroll
        for (int h = 0; h < 2; ++h) a[0] += step(h, 0);  // every near ray starts inside the source voxel
#pragma unroll
        for (int h = 0; h < 2; ++h) { const float dep = step(h, 1); a[1 + ns[h][1].loc] += dep; }
#pragma unroll
        for (int h = 0; h < 2; ++h) { const float dep = step(h, 2); a[7 + ns[h][2].loc] += dep; }
        // Every ray crosses shells 0-2; the steps past both rays' lengths (which would deposit 0)
        // are skipped by wave-uniform branches.
        const int lmax = max(len[0], len[1]);
        if (lmax > 3) {
            load_rho(4, 7);
#pragma unroll
            for (int h = 0; h < 2; ++h) { const float dep = step(h, 3); b[ns[h][3].loc] += dep; }
            if (lmax > 4) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {  // both places are updated, branch-free (x + 0 = x)
                    const float dep = step(h, 4);
                    const int loc = ns[h][4].loc;
                    const bool lo = loc < 32;
                    c[lo ? loc : 0] += lo ? dep : 0.0f;
                    a[lo ? 25 : loc - 7] += lo ? 0.0f : dep;
                }
                if (lmax > 5) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) { const float dep = step(h, 5); d[ns[h][5].loc] += dep; }
                    if (lmax > 6) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) { const float dep = step(h, 6); d[24 + ns[h][6].loc] += dep; }
                    }
                }
            }
        }
    }
}

// Bin q of the near block in the layout of near_rays.
#define NEAR_BIN(q)                                                                                      \
    (kShellMap.shell[q] == 0   ? na[0]                                                                  \
     : kShellMap.shell[q] == 1 ? na[1 + kShellMap.local[q]]                                             \
     : kShellMap.shell[q] == 2 ? na[7 + kShellMap.local[q]]                                             \
     : kShellMap.shell[q] == 3 ? nb[kShellMap.local[q]]                                                 \
     : kShellMap.shell[q] == 4 ? (kShellMap.local[q] < 32 ? nc[kShellMap.local[q]] : na[kShellMap.local[q] - 7]) \
     : kShellMap.shell[q] == 5 ? nd[kShellMap.local[q]]                                                 \
                               : nd[24 + kShellMap.local[q]])

// Near block over row segments: one wave = 64 consecutive sources of a row (density reads are
// uniform base + lane). Writes the 125 weights of every source to its own planes and to its
// targets' planes (bin q of source s is entry q of target s + d_q), and starts W with their sum.
template <bool EXACT, class WT, bool COUNT>
__global__ __launch_bounds__(256) void k_near(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                               int n_items, unsigned long long* vrs) {
    __shared__ float tab[kTabLds];
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    const int4 it = items[min(item_id, n_items - 1)];
    const Vol& V = vols[it.x];
    if (!EXACT) load_table(tab, V);
    if (item_id >= n_items || it.w <= 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int row = it.y, y = row % V.ny, z = row / V.ny, x0 = it.z;
    const bool valid = lane < it.w;
    const size_t loc = (size_t)row * V.nx + x0 + (valid ? lane : 0);
    const bool act = valid && gload(V.support + loc);
    const float* rho0 = V.rho_p + ((size_t)(z + 1) * (V.ny + 2) + (y + 1)) * (V.nx + 2) + x0 + 1;
    unsigned visits = 0;
    NEAR_ACCUMULATORS;
    near_rays<EXACT, COUNT>(V, tab, act, [&](int dp) { return gload(rho0 + dp + lane); }, NEAR_ARGS, visits);
    // Near weights on the scale of one far ray (prc.cu:205-218); W starts with their sum.
    const int R1 = V.n_rays + 1, nseg = V.nseg;
    const int seg = (x0 - gload(V.row_xa + row)) / kWave;
    const RowInfo ri = gload(V.rinfo + ((size_t)row * nseg + seg) * R1);
    WT* wp = (WT*)V.wbuf + ri.off;
    WT* tb = (WT*)V.tbuf;
    const float nscale = (float)V.n_rays / (float)kNearRays;
#pragma unroll
    for (int q = 0; q < kNQ; ++q) {
        const WT h = WeightIO<WT>::store_value(act ? NEAR_BIN(q) * nscale : 0.0f);
        if (valid) gstore(wp + (size_t)q * ri.len + lane, h);
    }
    // W and the target-major copies, from the source's own planes read back (each lane reads its
    // own stores) in a loop that is not unrolled, so that the bins above stay in registers.
    float wacc = 0.0f;
#pragma unroll 1
    for (int q = 0; q < kNQ; ++q) {
        const WT h = valid ? gload(wp + (size_t)q * ri.len + lane) : WeightIO<WT>::store_value(0.0f);
        const float wq = WeightIO<WT>::to_float(h);
        if (wq > 0.0f) wacc += wq;
        // A deposit implies an active target, inside its row's active range.
        const int ty = y + (q / 5 % 5 - 2), tz = z + (q / 25 - 2);
        if (__any(wq > 0.0f) && (unsigned)ty < (unsigned)V.ny && (unsigned)tz < (unsigned)V.nz) {
            const int t = tz * V.ny + ty;
            const TargetPair tp = target_pair(V.tinfo + (size_t)t * nseg * R1, R1, gload(V.row_xa + t),
                                              gload(V.row_len + t), x0 + (q % 5 - 2), q);
            long long idx;
            if (target_slot(tp, lane, idx) && wq > 0.0f) gstore(tb + idx, h);
        }
    }
    if (act) gstore(V.W + loc, wacc);
    if (COUNT && vrs) {
        const long long tot = wave_sum_ll((long long)visits);
        if (lane == 0 && tot) atomicAdd(vrs, (unsigned long long)tot);
    }
}

// Near block over a list of sources (voxel index, template id or -1), 64 arbitrary sources per
// wave with per-lane density addressing: the sources whose 5x5x5 neighbourhood is not one
// uniform density, plus one representative per uniform density, whose 125 weights and sum
// become that density's template (every source with the same uniform neighbourhood computes the
// same weights bit for bit, so the others copy them in k_near_fill).
template <bool EXACT, class WT, bool COUNT>
__global__ __launch_bounds__(256) void k_near_list(const Vol* __restrict__ vols, const int2* __restrict__ list,
                                                    int n_list, float* __restrict__ tmpl,
                                                    float* __restrict__ tmpl_sum, unsigned long long* vrs) {
    __shared__ float tab[kTabLds];
    const Vol& V = vols[0];
    if (!EXACT) load_table(tab, V);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n_list;
    const int2 e = list[act ? i : n_list - 1];
    const int nx = V.nx, ny = V.ny;
    const int v = e.x, x = v % nx, row = v / nx, y = row % ny, z = row / ny;
    const float* rb = V.rho_p + ((size_t)(z + 1) * (ny + 2) + (y + 1)) * (nx + 2) + x + 1;
    unsigned visits = 0;
    NEAR_ACCUMULATORS;
    near_rays<EXACT, COUNT>(V, tab, act, [&](int dp) { return gload(rb + dp); }, NEAR_ARGS, visits);
    if (COUNT && vrs) {
        const long long tot = wave_sum_ll((long long)visits);
        if ((threadIdx.x & (kWave - 1)) == 0 && tot) atomicAdd(vrs, (unsigned long long)tot);
    }
    if (!act) return;
    const int R1 = V.n_rays + 1, nseg = V.nseg;
    const RowInfo ri = gload(V.rinfo + ((size_t)row * nseg + (x - gload(V.row_xa + row)) / kWave) * R1);
    WT* wp = (WT*)V.wbuf + ri.off + (x - ri.xa);
    WT* tb = (WT*)V.tbuf;
    const float nscale = (float)V.n_rays / (float)kNearRays;
#pragma unroll
    for (int q = 0; q < kNQ; ++q) gstore(wp + (size_t)q * ri.len, WeightIO<WT>::store_value(NEAR_BIN(q) * nscale));
    // W, the template copy and the target-major copies, from the source's own planes read back
    // (each lane reads its own stores) in a loop that is not unrolled, so that the bins above
    // stay in registers.
    float wacc = 0.0f;
#pragma unroll 1
    for (int dzy = 0; dzy < 25; ++dzy) {
        WT h[5];
        float wq[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int q = dzy * 5 + k;
            h[k] = gload(wp + (size_t)q * ri.len);
            wq[k] = WeightIO<WT>::to_float(h[k]);
            if (wq[k] > 0.0f) wacc += wq[k];
            if (e.y >= 0) tmpl[e.y * kNQ + q] = wq[k];
        }
        // The five targets x-2 .. x+2 of row t lie in at most two of its segments.
        const int ty = y + dzy % 5 - 2, tz = z + dzy / 5 - 2;
        if ((unsigned)ty >= (unsigned)ny || (unsigned)tz >= (unsigned)V.nz) continue;
        const int t = tz * ny + ty, tlen = gload(V.row_len + t), txa = gload(V.row_xa + t);
        const int rel = x - 2 - txa;
        const int s0 = rel >= 0 ? rel / kWave : -((kWave - 1 - rel) / kWave);
        const RowInfo* u0 = V.tinfo + (size_t)t * nseg * R1;
        const int ns = (tlen + kWave - 1) / kWave;
        RowInfo t0{}, t1{};
        if (s0 >= 0 && s0 < ns) t0 = gload(u0 + (size_t)s0 * R1);
        if (s0 + 1 >= 0 && s0 + 1 < ns) t1 = gload(u0 + (size_t)(s0 + 1) * R1);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int q = dzy * 5 + dx + 2;
            // A deposit implies an active target: inside one of the two segments.
            const int tx = x + dx;
            const RowInfo& te = (t1.len > 0 && tx >= t1.xa) ? t1 : t0;
            if (wq[dx + 2] > 0.0f && tx >= te.xa && tx < te.xa + te.len && q < te.cnt)
                gstore(tb + te.off + (size_t)q * te.len + (tx - te.xa), h[dx + 2]);
        }
    }
    gstore(V.W + v, wacc);
    if (e.y >= 0) tmpl_sum[e.y] = wacc;
}

// Copy the templates to the sources with a uniform neighbourhood (cls >= 0) and write zeros for
// inactive voxels of the row segments (cls = -2); listed sources (cls = -1) were written by
// k_near_list. Same stores as k_near, values from the template of the source's density.
template <class WT>
__global__ __launch_bounds__(256) void k_near_fill(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                                    int n_items, const signed char* __restrict__ cls,
                                                    const float* __restrict__ tmpl,
                                                    const float* __restrict__ tmpl_sum) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    if (it.w <= 0) return;
    const Vol& V = vols[it.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int row = it.y, y = row % V.ny, z = row / V.ny, x0 = it.z;
    const bool valid = lane < it.w;
    const size_t loc = (size_t)row * V.nx + x0 + (valid ? lane : 0);
    const int c = valid ? (int)cls[loc] : -1;
    const bool write = valid && c != -1;
    const int R1 = V.n_rays + 1, nseg = V.nseg;
    const int seg = (x0 - gload(V.row_xa + row)) / kWave;
    const RowInfo ri = gload(V.rinfo + ((size_t)row * nseg + seg) * R1);
    WT* wp = (WT*)V.wbuf + ri.off;
    WT* tb = (WT*)V.tbuf;
    const float* tpl = tmpl + (c >= 0 ? c : 0) * kNQ;
    // The 125 bins by target row (dz, dy), whose records are read once: the five targets x0 - 2 ..
    // x0 + 2 (plus the lanes) fall into segments s_lo .. s_lo + 2 of the row, and a bin's first
    // segment is s_lo or s_lo + 1.
#pragma unroll 1
    for (int dzy = 0; dzy < 25; ++dzy) {
        const int ty = y + dzy % 5 - 2, tz = z + dzy / 5 - 2;
        const bool tin = (unsigned)ty < (unsigned)V.ny && (unsigned)tz < (unsigned)V.nz;
        int xa_t = 0, ns = 0, s_lo = 0;
        RowInfo e[3] = {};
        if (tin) {
            const int t = tz * V.ny + ty;
            xa_t = gload(V.row_xa + t);
            ns = (gload(V.row_len + t) + kWave - 1) / kWave;
            const int rel = x0 - 2 - xa_t;
            s_lo = rel >= 0 ? rel / kWave : -((kWave - 1 - rel) / kWave);  // floor
            const RowInfo* unit0 = V.tinfo + (size_t)t * nseg * R1;
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (s_lo + j >= 0 && s_lo + j < ns) e[j] = gload(unit0 + (size_t)(s_lo + j) * R1);
        }
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int q = dzy * 5 + dx + 2;
            const float w = c >= 0 ? tpl[q] : 0.0f;
            const WT h = WeightIO<WT>::store_value(w);  // exact: templates hold stored values
            if (write) gstore(wp + (size_t)q * ri.len + lane, h);
            if (tin && __any(w > 0.0f)) {
                // target_pair with the row's records at hand (near block: ray 0, step q).
                const int tx0 = x0 + dx, rel = tx0 - xa_t;
                const int s0 = rel >= 0 ? rel / kWave : -((kWave - 1 - rel) / kWave);  // s_lo or s_lo + 1
                TargetPair tp{0, 0, kWave * (s0 + 1) - rel, 0};
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    const int sg = s0 + hh;
                    if (sg >= 0 && sg < ns) {
                        const int j = sg - s_lo;  // fields picked by value: the records stay in registers
                        const long long off = j == 0 ? e[0].off : j == 1 ? e[1].off : e[2].off;
                        const int cnt = j == 0 ? e[0].cnt : j == 1 ? e[1].cnt : e[2].cnt;
                        const int xa = j == 0 ? e[0].xa : j == 1 ? e[1].xa : e[2].xa;
                        const int ln = j == 0 ? e[0].len : j == 1 ? e[1].len : e[2].len;
                        tp.ok |= q < cnt ? 1 << hh : 0;
                        const long long b = off - xa + (long long)q * ln + tx0;
                        if (hh) tp.b1 = b;
                        else tp.b0 = b;
                    }
                }
                long long idx;
                if (target_slot(tp, lane, idx) && w > 0.0f) gstore(tb + idx, h);
            }
        }
    }
    if (c >= 0) gstore(V.W + loc, tmpl_sum[c]);
}

// Near-block memo, classification. A source is uniform when all 125 densities of its 5x5x5
// block equal its own; a block reaching outside the grid includes the kHalo halo or margin and
// never is. Uniform densities get template ids from a 64-slot table (first come, first served:
// the ids only name templates, the weights do not depend on them), and the representative of
// each template is its lowest voxel index. cls: -2 inactive, -1 marched, t >= 0 template t.
constexpr int kMaxTemplates = 64;
constexpr unsigned kEmptyKey = 0xffffffffu;  // a NaN pattern; densities are finite

__device__ inline int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__global__ __launch_bounds__(256) void k_classify(const Vol* __restrict__ vols, signed char* __restrict__ cls,
                                                   unsigned* __restrict__ keys, int* __restrict__ rep) {
    const Vol& V = vols[0];
    const int nx = V.nx, ny = V.ny;
    const int n = nx * ny * V.nz;
    const long long vl = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const bool in = vl < n;
    const int v = in ? (int)vl : n - 1, vv = v;
    const int x = vv % nx, row = vv / nx, y = row % ny, z = row / ny;
    const bool act = in && V.support[vv];
    const long long sy = nx + 2, sz = (long long)(ny + 2) * (nx + 2);
    const float* c = V.rho_p + (z + 1) * sz + (y + 1) * sy + x + 1;
    const float r0 = *c;
    bool uni = act;
    for (int dz = -2; dz <= 2 && uni; ++dz)
        for (int dy = -2; dy <= 2 && uni; ++dy) {
            const float* rr = c + dz * sz + dy * sy;
            uni = rr[-2] == r0 && rr[-1] == r0 && rr[0] == r0 && rr[1] == r0 && rr[2] == r0;
        }
    const unsigned bits = __float_as_uint(r0);
    int id = -1;
    bool need = uni;
    unsigned long long pending = __ballot(need);
    while (pending) {  // one distinct density of the wave per pass
        const int leader = __ffsll((long long)pending) - 1;
        const unsigned key = __shfl(bits, leader);
        int slot = -1;
        if (lane == leader)
            for (int k = 0; k < kMaxTemplates; ++k) {
                unsigned cur = __hip_atomic_load(keys + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == kEmptyKey) cur = atomicCAS(keys + k, kEmptyKey, key), cur = cur == kEmptyKey ? key : cur;
                if (cur == key) { slot = k; break; }
            }
        slot = __shfl(slot, leader);
        const bool mine = need && bits == key;
        if (slot >= 0) {
            const int m = wave_min_i(mine ? v : INT_MAX);
            if (lane == leader && m < __hip_atomic_load(rep + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(rep + slot, m);
        }
        if (mine) id = slot, need = false;
        pending = __ballot(need);
    }
    if (in) cls[v] = act ? (signed char)(uni ? id : -1) : (signed char)-2;
}

// The march list of the memo: marched sources and template representatives, appended per block
// (one atomic per block of kMemoChunks x 256 voxels: the blocks' appends to one counter serialise,
// so fewer and larger blocks; the order only changes which lanes march together). Also counts the
// uniform sources.
constexpr int kMemoBlock = 256;
constexpr int kMemoChunks = 16;
constexpr int kMemoWaves = kMemoBlock / kWave;
__global__ __launch_bounds__(kMemoBlock) void k_memo_list(int n, const signed char* __restrict__ cls,
                                                          const int* __restrict__ rep, int2* __restrict__ list,
                                                          int* __restrict__ ctr) {
    __shared__ int wt[kMemoChunks * kMemoWaves], base;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const long long v0 = (long long)blockIdx.x * kMemoChunks * kMemoBlock + threadIdx.x;
    auto taken = [&](int j, int& v, int& c) {
        const long long vl = v0 + (long long)j * kMemoBlock;
        v = (int)min(vl, (long long)n - 1);
        c = vl < n ? (int)cls[v] : -2;
        return c == -1 || (c >= 0 && rep[c] == v);
    };
    // Pass 1: the taken and uniform voxels of each chunk, per wave.
    int us = 0;
    for (int j = 0; j < kMemoChunks; ++j) {
        int v, c;
        const unsigned long long b = __ballot(taken(j, v, c));
        us += __popcll(__ballot(c >= 0));
        if (lane == 0) wt[j * kMemoWaves + wv] = __popcll(b);
    }
    __shared__ int wu[kMemoWaves];
    if (lane == 0) wu[wv] = us;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0, u = 0;
        for (int k = 0; k < kMemoChunks * kMemoWaves; ++k) t += wt[k];
        for (int k = 0; k < kMemoWaves; ++k) u += wu[k];
        base = t ? atomicAdd(ctr, t) : 0;
        if (u) atomicAdd(ctr + 1, u);
    }
    __syncthreads();
    // Pass 2: each chunk's taken voxels in order, after those of the chunks and waves before it.
    int off = base;
    for (int k = 0; k < wv; ++k) off += wt[k];
    for (int j = 0; j < kMemoChunks; ++j) {
        int v, c;
        const bool take = taken(j, v, c);
        const unsigned long long b = __ballot(take);
        if (take) list[off + __popcll(b & ((1ull << lane) - 1))] = make_int2(v, c);
        if (j + 1 < kMemoChunks) {
            for (int k = wv; k < kMemoWaves; ++k) off += wt[j * kMemoWaves + k];
            for (int k = 0; k < wv; ++k) off += wt[(j + 1) * kMemoWaves + k];
        }
    }
}

// Far rays, count pass: the number of stored steps of every (segment, ray), i.e. one past the last
// step at which some source of the segment deposits, less kfirst. The count only sizes the layout:
// an over-count stores extra zero planes and changes no result. A step deposits (prc.cu:146-163)
// only if the lane is alive (range below the cutoff), inside the grid, in tissue and past kfirst;
// the count takes the last such step, an upper bound of the last deposit that needs neither the
// survival table nor its lookup. The march itself (range, air length, death) is the fill pass's.
// One wave per segment; step lengths and shifts come from scalar loads (the step table is read
// through the constant address space: it is not written while kernels run). COUNT adds the
// number of voxel visits of live lanes to *vrs.
// Zeroes store instructions [i0, i1) of a wave's share of a clear range: instruction i writes the
// 64 16-byte units base + 64 i + lane that lie below `units`.
__device__ inline void clear_units(v4i* __restrict__ clear, long long units, long long base, long long i0,
                                   long long i1, int lane) {
    for (long long i = i0; i < i1; ++i) {
        const long long e = base + i * kWave + lane;
        if (e < units) *(GLOBAL_AS v4i*)(clear + e) = v4i{0, 0, 0, 0};
    }
}

// Far rays, count pass. `clear` (units 16-byte units, or none) is a buffer this pass zeroes beside
// the march: a rebuild clears the target-major copy it is about to refill here, where the march
// leaves the memory bus idle, instead of in a memory fill of its own after the layout. Each wave
// zeroes an equal share, spread over its rays (waves without a segment zero theirs at once).
template <bool COUNT>
__global__ __launch_bounds__(256) void k_far_count(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                                    int n_items, const FarStep* __restrict__ fstep,
                                                    const int* __restrict__ kfirst, RowInfo* __restrict__ rinfo,
                                                    unsigned long long* vrs, v4i* __restrict__ clear, long long units) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const long long waves = (long long)gridDim.x * kItemsPerBlock;
    const long long share = (units + waves * kWave - 1) / (waves * kWave);  // store instructions per wave
    const long long cbase = (long long)item_id * share * kWave;
    if (it.w <= 0) {
        clear_units(clear, units, cbase, 0, share, threadIdx.x & (kWave - 1));
        return;
    }
    const Vol& V = vols[it.x];
    const int ny = V.ny, n_rays = V.n_rays, R1 = n_rays + 1, LF = V.LF;
    const float cutoff = V.cutoff, support_rho = V.support_rho, air_cap = V.air_cap;
    const int lane = threadIdx.x & (kWave - 1);
    const int row = it.y, y = row % ny, z = row / ny, x0 = it.z;
    const bool valid = lane < it.w;
    const bool act = valid && gload(V.support + (size_t)row * V.nx + x0 + lane);
    // Every lane loads every step of a group, in bounds by the allocation's margin (a group runs
    // only while a lane is alive).
    const RhoReader rho_at(V.rho_p, ((size_t)(z + 1) * (ny + 2) + (y + 1)) * (V.nx + 2) + x0 + 1 + lane);
    const size_t unit = (size_t)row * V.nseg + (x0 - gload(V.row_xa + row)) / kWave;
    unsigned visits = 0;
    // Outside the support's bounding box, on an axis along which the ray moves away from the box
    // (or not at all), a lane never reaches a support voxel again (the walk's voxel coordinates are
    // monotonic), so it cannot deposit: such lanes are ended at the end of a group, which ends the
    // wave's air tail early. Only the statistics pass marches every step.
    const int* box = V.box;
    const int xl = x0 + lane;
    for (int r = 0; r < n_rays; ++r) {
        clear_units(clear, units, cbase, share * r / n_rays, share * (r + 1) / n_rays, lane);
        const int kf = kfirst[1 + r];
        const FarStep* fs = fstep + (size_t)r * LF;
        float w = act ? 0.0f : cutoff, air = 0.0f;
        // The ray's direction signs: those of its last tabulated step's offsets.
        const v4i fe = *(const CONST_AS v4i*)(fs + V.L - 1);
        const int ex = (short)(fe[2] & 0xffff), ey = fe[2] >> 16, ez = (short)(fe[3] & 0xffff);
        // The lane's last step that may deposit, before or after kf: the wave's count is the part
        // past kf, the same as when only steps from kf on were considered.
        int last = -1;
        const unsigned* fa = V.fsoa + (size_t)r * 2 * LF;  // the ray's seg bits, then its soffs
        for (int s0 = 0; s0 < LF && __any(w < cutoff); s0 += kCountGroup) {
            float seg[kCountGroup], rv[kCountGroup];
            // The group's lengths and offsets: 16-byte scalar loads at immediate offsets from two
            // bases (s0 and LF are multiples of 16, so every load is aligned).
#pragma unroll
            for (int u4 = 0; u4 < kCountGroup; u4 += 4) {
                const v4i gs = *(const CONST_AS v4i*)(fa + s0 + u4), go = *(const CONST_AS v4i*)(fa + LF + s0 + u4);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    seg[u4 + e] = __int_as_float(gs[e]);  // (hipcc miscompiles __builtin_bit_cast of a vector element)
                    rv[u4 + e] = rho_at.step((unsigned)go[e]);
                }
            }
            int lg = -1;  // within the group: u is an inline constant of the select
#pragma unroll
            for (int u = 0; u < kCountGroup; ++u) {
                const float rho = rv[u];
                const bool tissue = rho >= support_rho, alive = w < cutoff;  // kHalo: not tissue
                if (COUNT) visits += (rho < INFINITY && alive) ? 1u : 0u;
                lg = (tissue && alive) ? u : lg;
                air = tissue ? 0.0f : air + seg[u];
                // Death in air; a step onto the kHalo halo kills too, as fminf(fmaf(NaN, seg, w),
                // cutoff) = cutoff.
                w = air > air_cap ? cutoff : fminf(fmaf(rho, seg[u], w), cutoff);
            }
            last = lg >= 0 ? s0 + lg : last;
            if (!COUNT) {
                const v4i fc = *(const CONST_AS v4i*)(fs + s0 + kCountGroup - 1);  // the group's last voxel
                const int cx = xl + (short)(fc[2] & 0xffff), cy = y + (fc[2] >> 16), cz = z + (short)(fc[3] & 0xffff);
                const bool away_yz = (cy < box[2] && ey <= 0) || (cy > box[3] && ey >= 0) ||
                                     (cz < box[4] && ez <= 0) || (cz > box[5] && ez >= 0);
                const bool away = away_yz || (cx < box[0] && ex <= 0) || (cx > box[1] && ex >= 0);
                w = away ? cutoff : w;
            }
        }
        // One wave per segment: its count is the segment's.
        const int kmax = wave_max(last + 1);
        if (lane == 0 && kmax > kf) rinfo[unit * R1 + 1 + r].cnt = kmax - kf;
    }
    if (COUNT && vrs) {
        const long long tot = wave_sum_ll((long long)visits);
        if (lane == 0 && tot) atomicAdd(vrs, (unsigned long long)tot);
    }
}

// Far rays, fill pass: writes the weights and finishes W (continuing from the near sum, in the
// reference's accumulation order). The step table is read 64 steps at a time, one step per lane;
// the same lane also resolves the step's target row (the target-major plane offset) and stages
// the step in LDS, from where every lane reads it, so a step costs one density load, two LDS
// broadcast reads and the two weight stores. COUNT adds the number of stored positive far
// weights to *ndep.
template <bool EXACT, class WT, bool COUNT>
__global__ __launch_bounds__(256) void k_far_fill(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                                   int n_items, unsigned long long* ndep) {
    __shared__ float tab[kTabLds];
    __shared__ FillStep staged[kItemsPerBlock][kWave];  // per wave: the current 64 steps
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    const int4 it = items[min(item_id, n_items - 1)];
    const Vol& V = vols[it.x];
    if (!EXACT) load_table(tab, V);
    if (item_id >= n_items || it.w <= 0) return;
    const MarchConst M(V, tab, std::integral_constant<bool, EXACT>());
    const int ny = V.ny, nz = V.nz, L = V.L, LF = V.LF, n_rays = V.n_rays, R1 = n_rays + 1;
    const size_t units = (size_t)ny * nz * V.nseg;
    const FarStep* fstep = V.fstep;
    const int* kfirst = V.kfirst;
    const RowInfo* rinfo = V.rinfo;
    WT* wbuf = (WT*)V.wbuf;
    WT* tbuf = (WT*)V.tbuf;
    const int lane = threadIdx.x & (kWave - 1);
    const int row = it.y, y = row % ny, z = row / ny, x0 = it.z;
    const bool valid = lane < it.w;
    const size_t loc = (size_t)row * V.nx + x0 + (valid ? lane : 0);
    const bool act = valid && gload(V.support + loc);
    // Every lane loads every step of a group, in bounds by the allocation's margin (a group runs
    // only while a lane is alive).
    const RhoReader rho_at(V.rho_p, ((size_t)(z + 1) * (ny + 2) + (y + 1)) * (V.nx + 2) + x0 + 1 + lane);
    const int nseg = V.nseg;
    const size_t unit = (size_t)row * nseg + (x0 - gload(V.row_xa + row)) / kWave;  // this segment
    const int* row_xa = V.row_xa;
    const int* row_len = V.row_len;
    unsigned visits = 0, deposits = 0;
    float wacc = act ? gload(V.W + loc) : 0.0f;
    const int len = it.w;  // plane width = the segment's width
    // Byte offset of the lane's entry in a plane of its own run. Lanes past the segment's width get
    // an offset past the end of any run, so the run descriptor's range check drops their stores.
    const int wlane = valid ? lane * (int)sizeof(WT) : 0x7ffffff0;

    for (int r = 0; r < n_rays; ++r) {
        const int kf = gload(kfirst + 1 + r);
        const FarStep* fs = fstep + (size_t)r * LF;
        const RowInfo ri = gload(rinfo + unit * R1 + 1 + r);
        const int cnt = ri.cnt;
        if (cnt <= 0) continue;  // no source of this segment deposits along this ray
        // March only up to the row's last stored step.
        const int send = kf + cnt;
        // The run's planes kf .. send - 1 through a descriptor whose base is plane 0's place: the
        // store of step su goes to scalar offset su * len entries. (The values are wave-uniform;
        // readfirstlane puts them in scalar registers, where the descriptor must be.)
        const unsigned long long wb = (unsigned long long)(wbuf + ri.off - (long long)kf * len);
        const unsigned long long wbu = ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(wb >> 32)) << 32) |
                                       (unsigned)__builtin_amdgcn_readfirstlane((int)wb);
        const __amdgpu_buffer_rsrc_t wrsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)wbu, (short)0, __builtin_amdgcn_readfirstlane((int)((size_t)send * len * sizeof(WT))), 0x00020000);
        // Seg and soff of steps s .. s + kFarGroup - 1 (scalar loads, one per step: a group starts
        // at any step, and wide scalar loads need their natural alignment; s + kFarGroup - 1 < LF,
        // the table's zero steps past L keeping the reads in bounds) and the lanes' densities there.
        auto load_group = [&](int s, float (&seg)[kFarGroup], float (&rv)[kFarGroup]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < kFarGroup; ++u) {
                const v2i f = *(const CONST_AS v2i*)(fs + s + u);
                seg[u] = __int_as_float(f[0]);
                rv[u] = rho_at.step((unsigned)f[1]);
            }
        };
        Lane st = M.start(act);
        // Steps 0 .. kf - 1 lie in the near block: marched for the state they leave, without deposits.
        for (int s0 = 0; s0 < kf; s0 += kFarGroup) {
            float seg[kFarGroup], rv[kFarGroup];
            load_group(s0, seg, rv);
#pragma unroll
            for (int u = 0; u < kFarGroup; ++u)
                if (s0 + u < kf) (void)march_step<EXACT>(st, rv[u], seg[u], false, M, tab, visits);
        }
        // Steps kf .. send - 1, in groups of kFarGroup: step kf + i is plane i of the run. A full
        // group (FULL) stores all of its steps; the last group of the run may hold steps past send,
        // marched without stores (past the row's last stored step no lane of the row deposits: that
        // is what the count pass bounded).
        auto far_group = [&](auto full_c, int c0, int j, int nst, const FillStep* st_w) __attribute__((always_inline)) {
            constexpr bool FULL = decltype(full_c)::value;
            float seg[kFarGroup], rv[kFarGroup];
            load_group(c0 + j, seg, rv);
            WT hv[kFarGroup];
#pragma unroll
            for (int u = 0; u < kFarGroup; ++u) {
                const float dep = march_step<EXACT>(st, rv[u], seg[u], true, M, tab, visits);
                hv[u] = WeightIO<WT>::store_value(dep);
                // Deposits are >= 0 (or -0): adding the ones that are not positive changes nothing.
                const float wv = WeightIO<WT>::to_float(hv[u]);
                wacc += wv;
                if (COUNT && wv > 0.0f) ++deposits;
            }
            // The group's stores go out after its last density wait, so no wait of this group
            // waits on a store.
#pragma unroll
            for (int u = 0; u < kFarGroup; ++u) {
                if (FULL || j + u < nst) {
                    buffer_store_weight(hv[u], wrsrc, wlane, (c0 + j + u) * len * (int)sizeof(WT));
                    // Target-major copy: step su of this source is step su of its target. A deposit
                    // implies an active target inside its row's active range.
                    const FillStep q = st_w[j + u];
                    const bool holds = (lane < q.lo_lim) | (lane >= q.hi_start);
                    const long long idx = q.b0 + lane + (lane >= q.thr ? q.delta : 0);
                    if (holds & (WeightIO<WT>::to_float(hv[u]) > 0.0f)) gstore(tbuf + idx, hv[u]);
                }
            }
        };
        int s = kf;  // steps before s are done
        bool going = __any(M.alive(st));
        for (int c0 = kf; c0 < send && going; c0 += kWave) {
            const int sl = c0 + lane;
            const FarStep fl = sl < L ? gload(fs + sl) : FarStep{0.0f, kGatherBias, 0, 0, 0, 0};
            // Target-major slots of step sl (its target row's one or two segments).
            TargetPair tl{0, 0, kWave, 0};
            if (sl < send) {
                const int ty = y + fl.dy, tz = z + fl.dz;
                if ((unsigned)ty < (unsigned)ny && (unsigned)tz < (unsigned)nz) {
                    const int t = tz * ny + ty;
                    tl = target_pair_run(V.tmaj + (size_t)(1 + r) * units + (size_t)t * nseg, gload(row_xa + t),
                                         gload(row_len + t), x0 + fl.dx, sl - kf);
                }
            }
            FillStep* const st_w = staged[wave_in_block()];
            st_w[lane] = fill_step(tl);  // read back below by every lane of this wave only
            const int nst = min(kWave, send - c0);
            for (int j = 0; j < nst; j += kFarGroup) {
                if (j + kFarGroup <= nst)
                    far_group(std::true_type(), c0, j, nst, st_w);
                else
                    far_group(std::false_type(), c0, j, nst, st_w);
                s = c0 + j + kFarGroup;
                if (!__any(M.alive(st))) {
                    going = false;
                    break;
                }
            }
        }
        for (; s < send; ++s) buffer_store_weight(WeightIO<WT>::store_value(0.0f), wrsrc, wlane, s * len * (int)sizeof(WT));
    }
    if (act) gstore(V.W + loc, wacc);
    if (COUNT && ndep) {
        const long long tot = wave_sum_ll((long long)deposits);
        if (lane == 0 && tot) atomicAdd(ndep, (unsigned long long)tot);
    }
}

// Far rays marched inside a pass, for projectors whose stored matrix does not fit in device memory
// (only the near block is stored then; prc.cu:411-418 likewise retraces when its cache does not
// fit). One wave per source segment marches every far ray as the fill pass does, until every lane
// of the wave is dead, and uses each deposit w (rounded to the weight precision, as a stored weight
// would be) at once:
//   kMarchNorm: W_s += w (W holds the near block's sum; the same additions in the same order as
//               the fill, so W is the stored form's bit for bit);
//   kMarchAdj:  out_s += (sum w in[s + D]) / W_s on active s, after the stored pass wrote the near
//               block's share;
//   kMarchFwd:  out[s + D] += w g_s by atomic addition, after the stored pass wrote the near
//               block's share (the order of the additions, and so the last bits, vary between runs).
// A deposit lands only on a tissue voxel inside the grid, so every gather and every atomic stays
// inside the volume without a margin.
enum MarchMode { kMarchNorm = 0, kMarchAdj = 1, kMarchFwd = 2 };

template <bool EXACT, class WT, class VT, int MODE>
__global__ __launch_bounds__(256) void k_far_march(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                                    int n_items, const VT* __restrict__ vin, float* __restrict__ out) {
    __shared__ float tab[kTabLds];
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    const int4 it = items[min(item_id, n_items - 1)];
    const Vol& V = vols[it.x];
    if (!EXACT) load_table(tab, V);
    if (item_id >= n_items || it.w <= 0) return;
    const MarchConst M(V, tab, std::integral_constant<bool, EXACT>());
    const int ny = V.ny, L = V.L, LF = V.LF, LP = V.LP, n_rays = V.n_rays;
    const int lane = threadIdx.x & (kWave - 1);
    const int row = it.y, y = row % ny, z = row / ny, x0 = it.z;
    const bool valid = lane < it.w;
    const size_t loc = (size_t)row * V.nx + x0 + (valid ? lane : 0);
    const bool act = valid && gload(V.support + loc);
    const float wsrc = act ? gload(V.W + loc) : 0.0f;
    float gs = 0.0f;  // the forward's f_s / W_s (already divided; 0 skips the source)
    if (MODE == kMarchFwd) gs = act ? (float)vin[V.vbase + loc] : 0.0f;
    // Sources that contribute nothing are dead from the start.
    const bool live = MODE == kMarchFwd ? gs > 0.0f : MODE == kMarchAdj ? act && wsrc > 0.0f : act;
    if (!__any(live)) return;
    const RhoReader rho_at(V.rho_p, ((size_t)(z + 1) * (ny + 2) + (y + 1)) * (V.nx + 2) + x0 + 1 + lane);
    const VT* vbase = vin + V.vbase;
    float* obase = out + V.vbase;
    float acc = MODE == kMarchNorm ? wsrc : 0.0f;
    unsigned visits = 0;
    for (int r = 0; r < n_rays; ++r) {
        const int kf = gload(V.kfirst + 1 + r);
        const FarStep* fs = V.fstep + (size_t)r * LF;
        const int* plin = V.plin + (size_t)(1 + r) * LP;
        Lane st = M.start(live);
        for (int s0 = 0; s0 < L && __any(M.alive(st)); s0 += kFarGroup) {
            float seg[kFarGroup], rv[kFarGroup];
#pragma unroll
            for (int u = 0; u < kFarGroup; ++u) {  // s0 + kFarGroup - 1 < LF: zero steps past L
                const v2i f = *(const CONST_AS v2i*)(fs + s0 + u);
                seg[u] = __int_as_float(f[0]);
                rv[u] = rho_at.step((unsigned)f[1]);
            }
#pragma unroll
            for (int u = 0; u < kFarGroup; ++u) {
                const int k = s0 + u;
                const float dep = march_step<EXACT>(st, rv[u], seg[u], k >= kf, M, tab, visits);
                const float w = WeightIO<WT>::to_float(WeightIO<WT>::store_value(dep));
                if (MODE == kMarchNorm) {
                    acc += w;
                } else if (w > 0.0f) {  // k < L here: a deposit lies inside the grid
                    const size_t t = loc + (long long)*(const CONST_AS int*)(plin + min(k, LP - 1));
                    if (MODE == kMarchAdj) acc = fmaf(w, (float)vbase[t], acc);
                    else unsafeAtomicAdd(obase + t, w * gs);
                }
            }
        }
    }
    if (MODE == kMarchNorm && act) gstore(V.W + loc, acc);
    if (MODE == kMarchAdj && act && wsrc > 0.0f) obase[loc] += acc / wsrc;
}

// Plane offsets of the segment layout, computed on the device after the counting pass: the near
// block = 125 planes for every non-empty segment; empty segments hold nothing. The runs of a
// segment follow each other without a gap, and every segment starts on a 64-entry boundary, so a
// full segment's planes are whole cache lines.
__device__ inline int segment_width(int len, int seg) { return min(kWave, max(0, len - kWave * seg)); }
__host__ __device__ inline long long run_size(int cnt, int w) { return (long long)cnt * w; }
__host__ __device__ inline long long align_up(long long v, long long a) { return (v + a - 1) / a * a; }
// Shift lists: each segment's list starts on a multiple of kListAlign entries, and kListSlack zero
// entries follow the last list (reads of a segment's last plane group may run past its end).
// With FP16 weights the shifts of every aligned block of 8 planes are stored permuted (list_slot):
// the projector passes read plane pairs, and a lane's four planes 2g + h (g = 0..3) of a block
// must be four consecutive entries.
constexpr int kListAlign = 8;
constexpr int kListSlack = 64;
__host__ __device__ inline long long list_slot(long long i, bool pairs) {
    return pairs ? (i & ~7LL) | ((i & 1) << 2) | ((i >> 1) & 3) : i;
}

// Positive stored weights (nonzero bit patterns: stored weights are >= 0) among n 32-bit words
// of a weight buffer, into *ctr.
template <class WT>
__global__ void k_count_nonzero(const unsigned* __restrict__ words, long long n, unsigned long long* ctr) {
    unsigned long long c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const unsigned v = words[i];
        c += sizeof(WT) == 2 ? (unsigned)((v & 0xffffu) != 0) + (unsigned)((v >> 16) != 0) : (unsigned)(v != 0);
    }
    const long long tot = wave_sum_ll((long long)c);
    if ((threadIdx.x & (kWave - 1)) == 0 && tot) atomicAdd(ctr, (unsigned long long)tot);
}

// Far counts of every segment cleared (the projector stores the near block only).
__global__ void k_clear_far_counts(RowInfo* rinfo, long long units, int R1) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < units * R1 && id % R1 != 0) rinfo[id].cnt = 0;
}

// Inclusive prefix sum of v over the wave's lanes.
__device__ inline int wave_incl_sum(int v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int x = __shfl_up(v, o);
        if (lane >= o) v += x;
    }
    return v;
}

// Runs of one segment from their counts (cnt[p], lane = ray within a chunk of 64): each run's
// offset relative to the segment and its record; returns the segment's entries and planes. The
// runs of a segment all have its width, so the offsets are the width times a prefix sum of the
// counts. One wave per segment.
template <class CountOf>
__device__ inline void layout_runs(const Vol& V, RowInfo* info, int unit, int w, int x0, CountOf&& count_of,
                                   long long& entries, int& planes) {
    const int R1 = V.n_rays + 1, lane = threadIdx.x & (kWave - 1);
    entries = 0, planes = 0;
    for (int p0 = 0; p0 < R1; p0 += kWave) {
        const int p = p0 + lane;
        const int cnt = p < R1 && w > 0 ? count_of(p) : 0;
        const int incl = wave_incl_sum(cnt);
        if (p < R1) {
            RowInfo* e = info + (size_t)unit * R1 + p;
            e->cnt = cnt;
            e->off = entries + (long long)(incl - cnt) * w;  // run_size(cnt, w) = cnt * w
            e->xa = x0;
            e->len = w;
        }
        const int tot = __shfl(incl, kWave - 1);
        entries += (long long)tot * w;
        planes += tot;
    }
}

// Per segment: the runs' offsets relative to the segment, its total entries (aligned) and its
// planes (the list length, aligned). One wave per segment, lanes over rays.
__global__ __launch_bounds__(256) void k_row_totals(Vol V, long long* unit_tot, long long* unit_planes, PlaneRun* runs,
                                                     int* cnt_t) {
    const int units = V.ny * V.nz * V.nseg, R1 = V.n_rays + 1;
    const int unit = blockIdx.x * (blockDim.x / kWave) + wave_in_block();
    if (unit >= units) return;
    const int row = unit / V.nseg, seg = unit % V.nseg;
    const int w = segment_width(gload(V.row_len + row), seg), x0 = gload(V.row_xa + row) + kWave * seg;
    long long entries;
    int planes;
    layout_runs(V, V.rinfo, unit, w, x0, [&](int p) {
        const int cnt = p == 0 ? kNQ : V.rinfo[(size_t)unit * R1 + p].cnt;
        cnt_t[(size_t)p * units + unit] = cnt;  // ray-major copy of the counts
        return cnt;
    }, entries, planes);
    if (w == 0)
        for (int p = threadIdx.x & (kWave - 1); p < R1; p += kWave) cnt_t[(size_t)p * units + unit] = 0;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        unit_tot[unit] = align_up(entries, kWave);
        unit_planes[unit] = align_up(planes, kListAlign);
        runs[unit].planes = planes;
    }
}

// Exclusive prefix sums of the per-segment entries and list lengths, in three launches: the sums of
// 256-segment tiles, their scan by one block, and the scan inside each tile from its base. Small
// blocks: these launches run beside the march kernels of other projectors and must fit the CUs
// those leave free.
constexpr int kScanTile = 256;

// Exclusive scan of (va, vb) over the block's kScanTile threads; returns the tile totals in ta, tb.
__device__ inline void block_scan2(long long va, long long vb, long long& ea, long long& eb, long long& ta,
                                   long long& tb) {
    __shared__ long long wa[kScanTile / kWave], wb[kScanTile / kWave];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    __syncthreads();  // a previous call's readers are done with wa, wb
    long long sa = va, sb = vb;  // inclusive scan within the wave
    for (int o = 1; o < kWave; o <<= 1) {
        const long long xa = __shfl_up(sa, o), xb = __shfl_up(sb, o);
        if (lane >= o) sa += xa, sb += xb;
    }
    if (lane == kWave - 1) wa[wv] = sa, wb[wv] = sb;
    __syncthreads();
    long long pa = 0, pb = 0;
    ta = tb = 0;
    for (int k = 0; k < kScanTile / kWave; ++k) {
        if (k < wv) pa += wa[k], pb += wb[k];
        ta += wa[k], tb += wb[k];
    }
    ea = pa + sa - va, eb = pb + sb - vb;
}

__global__ __launch_bounds__(kScanTile) void k_tile_sums(const long long* __restrict__ a, const long long* __restrict__ b,
                                                          int n, long long* __restrict__ tile_a,
                                                          long long* __restrict__ tile_b) {
    const int i = blockIdx.x * kScanTile + threadIdx.x;
    long long ea, eb, ta, tb;
    block_scan2(i < n ? a[i] : 0, i < n ? b[i] : 0, ea, eb, ta, tb);
    if (threadIdx.x == 0) tile_a[blockIdx.x] = ta, tile_b[blockIdx.x] = tb;
}

// One block: exclusive bases of the tiles, in place, kScanTile at a time; the grand totals.
__global__ __launch_bounds__(kScanTile) void k_scan_tiles(long long* __restrict__ tile_a, long long* __restrict__ tile_b,
                                                           int tiles, long long* __restrict__ total_a,
                                                           long long* __restrict__ total_b) {
    long long ca = 0, cb = 0;  // totals of the chunks before
    for (int base = 0; base < tiles; base += kScanTile) {
        const int t = base + (int)threadIdx.x;
        long long ea, eb, ta, tb;
        block_scan2(t < tiles ? tile_a[t] : 0, t < tiles ? tile_b[t] : 0, ea, eb, ta, tb);
        if (t < tiles) tile_a[t] = ca + ea, tile_b[t] = cb + eb;
        ca += ta, cb += tb;
    }
    if (threadIdx.x == 0) *total_a = ca, *total_b = cb;
}

__global__ __launch_bounds__(kScanTile) void k_scan_in_tiles(const long long* __restrict__ a,
                                                              const long long* __restrict__ b, int n,
                                                              const long long* __restrict__ tile_a,
                                                              const long long* __restrict__ tile_b,
                                                              long long* __restrict__ a_base,
                                                              long long* __restrict__ b_base) {
    const int i = blockIdx.x * kScanTile + threadIdx.x;
    long long ea, eb, ta, tb;
    block_scan2(i < n ? a[i] : 0, i < n ? b[i] : 0, ea, eb, ta, tb);
    if (i < n) a_base[i] = tile_a[blockIdx.x] + ea, b_base[i] = tile_b[blockIdx.x] + eb;
}

// Absolute offsets: every run moves by its segment's base; the segment's plane sequence and list.
__global__ void k_row_offsets(Vol V, const long long* __restrict__ unit_base, const long long* __restrict__ list_base,
                              PlaneRun* runs) {
    const int R1 = V.n_rays + 1;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)V.ny * V.nz * V.nseg * R1) return;
    const long long unit = id / R1;
    V.rinfo[id].off += unit_base[unit];
    if (id % R1 == 0) runs[unit].off = unit_base[unit], runs[unit].dbase = (int)list_base[unit];
}

// kglob[p]: one past the last stored step of ray p over all segments (0 if it stores none), as
// kfirst(p) + the largest count. Blocks take (ray, a chunk of kKglobChunks of the segments) and
// fold their maximum in with one atomicMax (kglob starts at zero).
constexpr int kKglobChunks = 8;
__global__ __launch_bounds__(256) void k_kglob(Vol V, const int* __restrict__ cnt_t, int* __restrict__ kglob) {
    __shared__ int sm[256];
    const int p = blockIdx.x / kKglobChunks, chunk = blockIdx.x % kKglobChunks, units = V.ny * V.nz * V.nseg;
    const int* cnt_p = cnt_t + (size_t)p * units;  // the ray's counts, one per segment (coalesced)
    const int u0 = (int)((long long)units * chunk / kKglobChunks), u1 = (int)((long long)units * (chunk + 1) / kKglobChunks);
    int m = 0;
    for (int u = u0 + threadIdx.x; u < u1; u += blockDim.x) m = max(m, cnt_p[u]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && sm[0] > 0) atomicMax(kglob + p, gload(V.kfirst + p) + sm[0]);
}

// Forward bound per (segment, ray): one past the last step at which some source segment that
// stores that step reaches this segment's x range (an upper bound of where deposits land).
// Threads take (ray, segment) with the segment fastest, so neighbouring threads share the ray's
// steps, read neighbouring entries of the ray-major counts and write neighbouring entries of the
// ray-major bounds kt_t (which k_target_totals turns into the target-major runs).
__global__ void k_forward_bounds(Vol V, const int* __restrict__ kglob, const int* __restrict__ cnt_t,
                                 int* __restrict__ kt_t) {
    const int R1 = V.n_rays + 1, nseg = V.nseg, units = V.ny * V.nz * nseg;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)units * R1) return;
    const int unit = (int)(id % units), p = (int)(id / units);
    const int row = unit / nseg, seg = unit % nseg;
    const int tw = segment_width(gload(V.row_len + row), seg);
    int kt = 0;
    if (tw > 0 && p > 0) {  // the near block (p = 0) always holds its 125 bins
        const int tx0 = gload(V.row_xa + row) + kWave * seg, y = row % V.ny, z = row / V.ny;
        const int kf = gload(V.kfirst + p), kend = gload(kglob + p);
        const int* cnt_p = cnt_t + (size_t)p * units;
        // Downward from the ray's global bound: the first step that reaches this segment is the bound.
        // Four steps at a time, so that their dependent loads (step, source row, counts) overlap.
        // The sources that reach x range [tx0, tx0 + tw) of this row at step k lie in x range
        // [tx0 - dx, tx0 - dx + tw) of row (y - dy, z - dz): at most two of its segments.
        constexpr int U = 4;
        for (int k = kend - 1; k >= kf && kt == 0; k -= U) {
            int4 d[U];
            int srow[U], slen[U], sxa[U], c0[U], c1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) d[u] = k - u >= kf ? gload(V.pstep + (size_t)p * V.LP + k - u) : int4{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int ys = y - d[u].y, zs = z - d[u].z;
                const bool in = k - u >= kf && (unsigned)ys < (unsigned)V.ny && (unsigned)zs < (unsigned)V.nz;
                srow[u] = in ? zs * V.ny + ys : -1;
                slen[u] = in ? gload(V.row_len + srow[u]) : 0;
                sxa[u] = in ? gload(V.row_xa + srow[u]) : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int lo = max(tx0 - d[u].x, sxa[u]), hi = min(tx0 - d[u].x + tw, sxa[u] + slen[u]);
                c0[u] = c1[u] = 0;
                if (lo < hi) {
                    const int s0 = (lo - sxa[u]) / kWave, s1 = (hi - 1 - sxa[u]) / kWave;
                    c0[u] = gload(cnt_p + (size_t)srow[u] * nseg + s0);
                    if (s1 != s0) c1[u] = gload(cnt_p + (size_t)srow[u] * nseg + s1);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (kt == 0 && (k - u - kf < c0[u] || k - u - kf < c1[u])) kt = k - u + 1;
        }
    }
    kt_t[id] = kt;  // ray-major, as the threads run: [p][unit]
}

// Target-major layout: segment t holds, for ray p, the steps kfirst(p) .. kt(t, p) - 1 at which some
// source deposits into it (the near block: all 125 bins); planes over the segment's x range, runs
// without gaps as in the source-major copy.
__global__ __launch_bounds__(256) void k_target_totals(Vol V, long long* unit_tot, long long* unit_planes,
                                                        PlaneRun* runs, const int* __restrict__ kt_t) {
    const int units = V.ny * V.nz * V.nseg;
    const int unit = blockIdx.x * (blockDim.x / kWave) + wave_in_block();
    if (unit >= units) return;
    const int row = unit / V.nseg, seg = unit % V.nseg;
    const int w = segment_width(gload(V.row_len + row), seg), x0 = gload(V.row_xa + row) + kWave * seg;
    long long entries;
    int planes;
    layout_runs(V, V.tinfo, unit, w, x0, [&](int p) {
        return p == 0 ? kNQ : max(0, kt_t[(size_t)p * units + unit] - gload(V.kfirst + p));
    }, entries, planes);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        unit_tot[unit] = align_up(entries, kWave);
        unit_planes[unit] = align_up(planes, kListAlign);
        runs[unit].planes = planes;
    }
}

__global__ void k_target_offsets(Vol V, const long long* __restrict__ unit_base,
                                 const long long* __restrict__ list_base, PlaneRun* runs, TargetRun* tmaj) {
    const int R1 = V.n_rays + 1;
    const long long units = (long long)V.ny * V.nz * V.nseg;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= units * R1) return;
    const long long unit = id / R1;
    RowInfo& e = V.tinfo[id];
    e.off += unit_base[unit];
    tmaj[(id % R1) * units + unit] = TargetRun{e.off, e.cnt, 0};
    if (id % R1 == 0) runs[unit].off = unit_base[unit], runs[unit].dbase = (int)list_base[unit];
}

// Shift lists: the entry of step k of ray p's run is sign * D(p, kfirst(p) + k), segment by segment
// in plane order; the alignment padding of each list is zero. One wave per segment: the wave
// reads the runs' counts 64 rays at a time and places each run by a prefix sum over its lanes.
__global__ __launch_bounds__(256) void k_plane_lists(Vol V, const RowInfo* __restrict__ info,
                                                      const PlaneRun* __restrict__ runs, int sign, int pairs,
                                                      int* __restrict__ list) {
    const int units = V.ny * V.nz * V.nseg, R1 = V.n_rays + 1;
    const int unit = blockIdx.x * (blockDim.x / kWave) + wave_in_block();
    if (unit >= units) return;
    const int lane = threadIdx.x & (kWave - 1);
    const PlaneRun r = runs[unit];
    if (r.planes == 0) return;
    int* out = list + r.dbase;
    int pos = 0;  // planes of the rays before this chunk
    for (int p0 = 0; p0 < R1; p0 += kWave) {
        const int p = p0 + lane;
        const int cnt = p < R1 ? info[(size_t)unit * R1 + p].cnt : 0;
        const int kf = p < R1 ? gload(V.kfirst + p) : 0;
        int incl = cnt;  // inclusive prefix of the counts over the lanes
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int x = __shfl_up(incl, o);
            if (lane >= o) incl += x;
        }
        const int frel = incl - cnt;  // the lane's run starts at plane frel of the chunk
        const int tot = __shfl(incl, kWave - 1);
        // Lanes take the chunk's planes in stripes of 64 (coalesced stores); the run holding plane q
        // is the last whose start is at most q (runs of zero planes share the next run's start and
        // lose to it), found by a binary search over the lanes' starts.
        // (The loop runs the same number of times on every lane, so that every shuffle reads
        // lanes that take part in it.)
        for (int q0 = 0; q0 < tot; q0 += kWave) {
            const int q = q0 + lane;
            int j = 0, fj = 0;
#pragma unroll
            for (int step = kWave / 2; step >= 1; step >>= 1) {
                const int fc = __shfl(frel, j + step);
                if (fc <= q) j += step, fj = fc;
            }
            const int k0 = __shfl(kf, j);
            if (q < tot) out[list_slot(pos + q, pairs)] = sign * gload(V.plin + (size_t)(p0 + j) * V.LP + k0 + (q - fj));
        }
        pos += tot;
    }
    if (lane < align_up(pos, kListAlign) - pos) out[list_slot(pos + lane, pairs)] = 0;
}

// ---------------------------------------------------------------------------------------
// Projector passes.

// g = f / W on active sources with f / W > 0 (ForwardOp::begin, prc.cu:244-248), else 0.
__device__ inline float forward_scale(float f, float w) {
    const float s = w > 0.0f ? f / w : 0.0f;
    return s > 0.0f ? s : 0.0f;
}

// Plane-parallel streaming. A wave streams the plane sequence of one segment. Each lane loads 16
// bytes of a plane: E = 16 / sizeof(WT) consecutive entries (x = c*E .. c*E + E - 1 with
// c = lane % LPP), so one load instruction covers PPI = E planes, and a group of G loads covers
// P = PPI * G = 32 planes. With FP32 weights lane sub-plane s = lane / LPP takes planes
// q + s*G + g (g < G) of the group starting at plane q: each load reads 4 planes of 256 bytes.
// With FP16 weights (128-byte planes) load g reads the plane pairs 8j + 2g, 8j + 2g + 1 (j < 4),
// which lie next to each other: 4 runs of 256 bytes, which the texture units read in about 3/4 of
// the time of 8 separate planes; lane sub-plane s = 2j + h takes planes 8j + 2g + h. Either way a
// lane's G shifts are consecutive list entries (16-byte loads; FP16 lists are stored permuted
// within blocks of 8, see list_slot). Each lane keeps E running sums; at the end the PPI lanes of
// each x chunk are added together.
template <class WT>
struct PlaneLanes {
    static constexpr int E = 16 / (int)sizeof(WT);  // entries per lane and load
    static constexpr int LPP = kWave / E;           // lanes per plane
    static constexpr int PPI = E;                   // planes per load instruction
    static constexpr int G = 32 / PPI;              // loads per group
    static constexpr int P = PPI * G;               // planes per group
    static_assert(G == 4 || G == 8, "a lane's shifts load as int4s");
    static_assert(kListAlign % 8 == 0, "lists start on int4 boundaries and hold whole 8-plane blocks");
    static_assert(2 * P <= kListSlack, "the shift prefetch of a last group reads up to 2P - 1 past a list");
};

// The streaming loads are buffer loads: a wave-uniform descriptor and a 32-bit per-lane byte
// offset. Gathers address the vector through a base moved back by kGatherBias bytes, so that the
// lane's offset (bias + (x + shift) * element size) is nonnegative for shifts of either sign;
// shifts stay below 2^29 voxels in magnitude (a volume's density fits in 2^31 bytes), so the
// offset fits in 32 bits.

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

// N consecutive values of T (float or __half) at byte offset `off` of the buffer, as floats (8, 16
// or 32 bytes; the offset need only be aligned to T: the device reads them unaligned). POL is the
// loads' cache-policy field (1: sc0, which serves the load from L2 without allocating it in the
// CU's L1; the tuner picks it per pass).
template <class T, int N, int POL>
__device__ inline void load_gather(__amdgpu_buffer_rsrc_t r, unsigned off, float (&out)[N]) {
    constexpr int bytes = N * (int)sizeof(T);
    if constexpr (bytes % 16 == 0) {
#pragma unroll
        for (int h = 0; h < bytes / 16; ++h) {
            const v4i x = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(off + 16 * h), 0, POL));
            if constexpr (sizeof(T) == 2) {
                const h8 v = __builtin_bit_cast(h8, x);
#pragma unroll
                for (int e = 0; e < 8; ++e) out[8 * h + e] = (float)v[e];
            } else {
                const f4 v = __builtin_bit_cast(f4, x);
#pragma unroll
                for (int e = 0; e < 4; ++e) out[4 * h + e] = v[e];
            }
        }
    } else {
        static_assert(bytes == 8, "loads of 8 or a multiple of 16 bytes");
        const v2i x = __builtin_bit_cast(v2i, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, POL));
        if constexpr (sizeof(T) == 2) {
            const h4 v = __builtin_bit_cast(h4, x);
#pragma unroll
            for (int e = 0; e < 4; ++e) out[e] = (float)v[e];
        } else {
            const f2 v = __builtin_bit_cast(f2, x);
            out[0] = v[0], out[1] = v[1];
        }
    }
}

// E weights (16 bytes) at byte offset `off` of the buffer, as floats, or zeros unless `keep` (the
// zeroing is done on the loaded bits, so that each weight stays a plain conversion that the FMA
// consuming it can absorb); sc0 set (see segment_planes).
template <class WT, int E>
__device__ inline void load_weights(__amdgpu_buffer_rsrc_t r, unsigned off, float (&out)[E], bool keep) {
    static_assert(E * sizeof(WT) == 16, "one 16-byte load per lane");
    v4i x = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 1));
    if (!keep) x = v4i{0, 0, 0, 0};
    if constexpr (sizeof(WT) == 2) {
        const h8 h = __builtin_bit_cast(h8, x);
#pragma unroll
        for (int e = 0; e < E; ++e) out[e] = (float)h[e];
    } else {
        const f4 f = __builtin_bit_cast(f4, x);
#pragma unroll
        for (int e = 0; e < E; ++e) out[e] = f[e];
    }
}

template <int G>
__device__ inline void load_shifts(const GLOBAL_AS int* p, int (&d)[G]) {
#pragma unroll
    for (int h = 0; h < G; h += 4) {
        const v4i x = *(const GLOBAL_AS v4i*)(p + h);
        d[h] = x[0], d[h + 1] = x[1], d[h + 2] = x[2], d[h + 3] = x[3];
    }
}

// acc[e] += sum over the segment's planes k of w[k][x] * vec0[x + shift(k)] for the lane's x =
// c*E + e, all lanes ending with the full sums. `wseg` is the segment's first plane, `len` its
// width (the plane stride), `list` its shifts. Zero weights (no source) are skipped by a select
// where the gathered vector is FP32 (see below for the FP16 form).
// The last group clamps its plane reads to the last plane and zeroes the weights past it. Each
// group loads the next group's shifts while its own loads are in flight, so shift reads run up to
// 2P - 1 entries past the list (into the next list or the slack: valid shifts either way).
// Lanes with x >= len read the next plane's entries; their sums are not used.
template <class WT, class VT, int GPOL>
__device__ __forceinline__ void segment_planes(const WT* wseg, int len, int planes, const int* list, const VT* vec0,
                                      int lane, float (&acc)[PlaneLanes<WT>::E]) {
    using L = PlaneLanes<WT>;
    constexpr int E = L::E, LPP = L::LPP, G = L::G, P = L::P;
    const int s = lane / LPP, c = lane % LPP;
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = 0.0f;
    // The weight planes are read through a buffer descriptor over the segment's sequence (32-bit
    // offsets: a segment's planes stay far below 2 GB), with cache-policy bit sc0 set on the loads;
    // both were measured faster on the weight stream than plain global loads, the bit by about 5 %.
    const __amdgpu_buffer_rsrc_t wrsrc = buffer_of(wseg, 0x7fffffffLL);
    unsigned wq = 0;  // byte offset of the current group's first plane
    const unsigned group_bytes = (unsigned)(P * len * (int)sizeof(WT));
    // Plane of the group read by this lane's load g (see above).
    auto plane_of = [&](int g) { return sizeof(WT) == 2 ? 8 * (s >> 1) + 2 * g + (s & 1) : s * G + g; };
    unsigned woff[G];
#pragma unroll
    for (int g = 0; g < G; ++g) woff[g] = (unsigned)((plane_of(g) * len + c * E) * (int)sizeof(WT));
    // Gathers through a buffer descriptor whose base is moved back by kGatherBias bytes (only
    // base + offset is ever dereferenced, always inside the vector's allocation).
    const __amdgpu_buffer_rsrc_t grsrc =
        __builtin_amdgcn_make_buffer_rsrc((void*)((const char*)vec0 - kGatherBias), (short)0, (int)0xffffffffu, 0x00020000);
    const unsigned vlane = kGatherBias + (unsigned)(c * E * (int)sizeof(VT));
    const GLOBAL_AS int* lp = (const GLOBAL_AS int*)list + s * G;
    int d[G], dn[G];  // this lane's G shifts of the current and of the next group
    load_shifts<G>(lp, d);
    for (int q = 0; q < planes; q += P, wq += group_bytes) {
        lp += P;
        load_shifts<G>(lp, dn);
        const bool full = q + P <= planes;
        float w[G][E], v[G][E];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            unsigned off = woff[g];
            if (!full) off = (unsigned)((min(plane_of(g), planes - 1 - q) * len + c * E) * (int)sizeof(WT));
            load_weights<WT, E>(wrsrc, wq + off, w[g], full || q + plane_of(g) < planes);
            load_gather<VT, E, GPOL>(grsrc, vlane + (unsigned)d[g] * (unsigned)sizeof(VT), v[g]);
        }
        // Zero weights must not carry a non-finite input value into the sum (float vectors come
        // from callers of the operators). FP16 copies are finite by construction (written in
        // [0, 65504] by saturation, margins zero), and for finite v, fmaf(0, v, acc) is acc bit for
        // bit (acc starts at +0 and never becomes -0), so they need no select: one mixed-precision
        // FMA per entry.
        constexpr bool kFinite = std::is_same<VT, __half>::value;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int e = 0; e < E; ++e)
                acc[e] = (kFinite || w[g][e] != 0.0f) ? fmaf(w[g][e], v[g][e], acc[e]) : acc[e];
#pragma unroll
        for (int g = 0; g < G; ++g) d[g] = dn[g];
    }
#pragma unroll
    for (int o = LPP; o < kWave; o <<= 1)
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] += __shfl_xor(acc[e], o);
}

// Work distribution of the projector passes. The hardware places block b on XCD b % 8, and each
// XCD has its own L2. The item list (rows in z-major order) is cut into one contiguous chunk per
// XCD; a wave takes items from its own XCD's chunk, so the gathered neighbourhoods of the rows an
// XCD works on stay in that XCD's L2, and moves on to the next chunks once its own is drained
// (load balance). The counters are cleared by k_reset_queue before every launch.
constexpr int kXcds = 8;

__global__ void k_reset_queue(int* queue) {
    if (threadIdx.x < kXcds) queue[threadIdx.x] = 0;
}

__device__ inline int next_item(int* queue, int n_items, int xcd, int& slot) {
    const int lane = threadIdx.x & (kWave - 1);
    while (slot < kXcds) {
        const int c = (xcd + slot) % kXcds;
        const int b = (int)((long long)n_items * c / kXcds), e = (int)((long long)n_items * (c + 1) / kXcds);
        int i = 0;
        if (lane == 0) i = atomicAdd(queue + c, 1);
        i = __builtin_amdgcn_readfirstlane(__shfl(i, 0));
        if (b + i < e) return b + i;
        ++slot;
    }
    return -1;
}

// Schedules of the projector passes: one wave per item in launch order (grid), XCD-local work
// queues (queue), or launch order with each XCD's blocks renumbered to one contiguous chunk of
// the item list (chunked: block b runs on XCD b % 8; the j-th block of XCD x takes the j-th
// block-sized piece of chunk x), so that an XCD's gathers stay in its own L2 without queues.
enum Sched { kSchedGrid = 0, kSchedQueue = 1, kSchedChunked = 2 };
constexpr int kSchedCount = 3;

// A pass's launch shape, chosen per solver or batch by tune_passes: the schedule (enum Sched), for
// the queue schedule its persistent grid in blocks per CU, and the gathers' cache policy (0, or 1:
// sc0, served from L2 without allocating in the CU's L1).
struct PassShape {
    int sched = kSchedQueue;
    int blocks_per_cu = 2;
    int gpol = 0;
};
const PassShape kShapes[] = {
    {kSchedGrid, 0, 0},  {kSchedChunked, 0, 0}, {kSchedQueue, 1, 0}, {kSchedQueue, 2, 0}, {kSchedQueue, 4, 0},
    {kSchedGrid, 0, 1},  {kSchedChunked, 0, 1}, {kSchedQueue, 1, 1}, {kSchedQueue, 2, 1}, {kSchedQueue, 4, 1},
};
constexpr int kShapeCount = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
constexpr int kDefaultShape = 3;  // queue, two blocks per CU, gathers through L1
std::string shape_name(const PassShape& p) {
    static const char* const sched[kSchedCount] = {"grid", "queue", "chunked"};
    std::string n = sched[p.sched];
    if (p.sched == kSchedQueue) n += "@" + std::to_string(p.blocks_per_cu);
    if (p.gpol) n += "+sc0";
    return n;
}


__device__ inline unsigned chunked_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x, x = b % kXcds, j = b / kXcds;
    const unsigned q = nb / kXcds, r = nb % kXcds;  // XCD x runs q + (x < r) blocks
    return x * q + min(x, r) + j;
}

// out = A g where g already holds f / W (one item; see k_forward).
template <class WT, class VT, int GPOL>
__device__ __forceinline__ void forward_item(const Vol& V, const int4 it, const VT* __restrict__ gin,
                                    float* __restrict__ out, int lane) {
    constexpr int E = PlaneLanes<WT>::E, LPP = PlaneLanes<WT>::LPP;
    const int row = it.y, x0 = it.z;
    float* o = out + V.vbase + (size_t)row * V.nx + x0;
    if (it.w < 0) {  // voxels outside the row's active range receive nothing (prc.cu:432 clears out)
        if (lane < -it.w) o[lane] = 0.0f;
        return;
    }
    const size_t unit = (size_t)row * V.nseg + (x0 - gload(V.row_xa + row)) / kWave;
    const PlaneRun r = gload(V.truns + unit);
    float acc[E];
    segment_planes<WT, VT, GPOL>((const WT*)V.tbuf + r.off, it.w, r.planes, V.tlist + r.dbase,
                           gin + V.vbase + (size_t)row * V.nx + x0, lane, acc);
    if (lane < LPP)
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (lane * E + e < it.w) o[lane * E + e] = acc[e];
}

// out = A g where g already holds f / W (prc.cu:239-254 as a gather). The target's planes are
// streamed; for step (p, k) target t gathers g at t - D(p,k), a shift shared by the plane. The
// input vector carries a margin wide enough for every stored shift, so no gather needs a bounds
// check; zero weights (no source) are skipped by a select where the vector is FP32 (an FP16 vector
// is finite). Items are the active segments (w > 0)
// and the rest of every row (w < 0, written as zeros), so every voxel of the volume is written.
// SCHED selects the schedule (enum Sched), GPOL the gathers' cache policy. The gathered vector is
// FP32, or FP16 in the mixed form (VT).
template <class WT, class VT, int SCHED, int GPOL>
__global__ __launch_bounds__(256) void k_forward(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                                  int n_items, const VT* __restrict__ gin,
                                                  float* __restrict__ out, int* __restrict__ queue) {
    const int lane = threadIdx.x & (kWave - 1);
    if (SCHED != kSchedQueue) {
        const unsigned blk = SCHED == kSchedChunked ? chunked_block() : blockIdx.x;
        const int i = (int)blk * kItemsPerBlock + wave_in_block();
        if (i >= n_items) return;
        const int4 it = items[i];
        if (it.w != 0) forward_item<WT, VT, GPOL>(vols[it.x], it, gin, out, lane);
        return;
    }
    const int xcd = blockIdx.x % kXcds;
    int slot = 0;
    for (int i; (i = next_item(queue, n_items, xcd, slot)) >= 0;) {
        const int4 it = items[i];
        if (it.w != 0) forward_item<WT, VT, GPOL>(vols[it.x], it, gin, out, lane);
    }
}

// out_s = (sum_t w_st in_t) / W_s on active s (one item; see k_adjoint).
template <class WT, class VT, int GPOL>
__device__ __forceinline__ void adjoint_item(const Vol& V, const int4 it, const VT* __restrict__ in,
                                    float* __restrict__ out, int lane) {
    constexpr int E = PlaneLanes<WT>::E, LPP = PlaneLanes<WT>::LPP;
    const int row = it.y, x0 = it.z;
    const size_t unit = (size_t)row * V.nseg + (x0 - gload(V.row_xa + row)) / kWave;
    const PlaneRun r = gload(V.sruns + unit);
    float acc[E];
    segment_planes<WT, VT, GPOL>((const WT*)V.wbuf + r.off, it.w, r.planes, V.slist + r.dbase,
                           in + V.vbase + (size_t)row * V.nx + x0, lane, acc);
    if (lane < LPP)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int x = lane * E + e;
            const size_t loc = (size_t)row * V.nx + x0 + x;
            if (x < it.w && gload(V.support + loc)) {
                const float wv = gload(V.W + loc);
                out[V.vbase + loc] = wv > 0.0f ? acc[e] / wv : 0.0f;
            }
        }
}

// out_s = (sum_t w_st in_t) / W_s on active s (AdjointOp, prc.cu:256-267): stream the source's
// planes and gather in[] at s + D(p,k) (one shift per plane, margin-padded input).
template <class WT, class VT, int SCHED, int GPOL>
__global__ __launch_bounds__(256) void k_adjoint(const Vol* __restrict__ vols, const int4* __restrict__ items,
                                                  int n_items, const VT* __restrict__ in,
                                                  float* __restrict__ out, int* __restrict__ queue) {
    const int lane = threadIdx.x & (kWave - 1);
    if (SCHED != kSchedQueue) {
        const unsigned blk = SCHED == kSchedChunked ? chunked_block() : blockIdx.x;
        const int i = (int)blk * kItemsPerBlock + wave_in_block();
        if (i >= n_items) return;
        const int4 it = items[i];
        if (it.w > 0) adjoint_item<WT, VT, GPOL>(vols[it.x], it, in, out, lane);
        return;
    }
    const int xcd = blockIdx.x % kXcds;
    int slot = 0;
    for (int i; (i = next_item(queue, n_items, xcd, slot)) >= 0;) {
        const int4 it = items[i];
        if (it.w > 0) adjoint_item<WT, VT, GPOL>(vols[it.x], it, in, out, lane);
    }
}

// ---------------------------------------------------------------------------------------
// Image updates (prc.cu:269-332), one thread per voxel over row-segment items.

#define ITEM_PROLOGUE                                                          \
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();    \
    if (item_id >= n_items) return;                                            \
    const int4 it = items[item_id];                                            \
    const int lane = threadIdx.x & (kWave - 1);                                \
    if (lane >= it.w) return;                                                  \
    const Vol& V = vols[it.x];                                                 \
    const size_t loc = (size_t)it.y * V.nx + it.z + lane;                      \
    const size_t i = V.vbase + loc;                                            \
    (void)i;

// Largest value of v over the wave's lanes.
__device__ inline float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}

// Per-item partial maximum of the values a kernel writes (mixed form), for k_reduce_volume.
#define ITEM_MAX(v)                                                                              \
    if (pmax) {                                                                                  \
        const float m_ = wave_max_f(lane < it.w ? (v) : 0.0f);                                   \
        if (lane == 0) pmax[item_id] = (double)m_;                                               \
    }

// g = f / W for the forward (ForwardOp::begin). pmax, when given, receives per-item maxima of g
// (the mixed form scales its FP16 copy by them).
__global__ void k_fscale(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                         const float* __restrict__ f, float* __restrict__ g, double* __restrict__ pmax) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const int lane = threadIdx.x & (kWave - 1);
    const Vol& V = vols[it.x];
    float v = 0.0f;
    if (lane < it.w) {
        const size_t loc = (size_t)it.y * V.nx + it.z + lane, i = V.vbase + loc;
        v = gload(V.support + loc) ? forward_scale(f[i], gload(V.W + loc)) : 0.0f;
        g[i] = v;
    }
    ITEM_MAX(v)
}

// Biggs-Andrews prediction y = max(f + lambda (f - prev), f / 2) (prc.cu:318-321), fused with
// the forward scaling of y.
__global__ void k_extrapolate(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                              const Scal* __restrict__ sc, const float* __restrict__ f,
                              const float* __restrict__ prev, float* __restrict__ y, float* __restrict__ g,
                              double* __restrict__ pmax) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const int lane = threadIdx.x & (kWave - 1);
    const Vol& V = vols[it.x];
    float v = 0.0f;
    if (lane < it.w) {
        const size_t loc = (size_t)it.y * V.nx + it.z + lane, i = V.vbase + loc;
        const float lambda = sc[it.x].lambda;
        const float fi = f[i];
        const float yi = fmaxf(fmaf(lambda, fi - prev[i], fi), 0.5f * fi);
        y[i] = yi;
        v = gload(V.support + loc) ? forward_scale(yi, gload(V.W + loc)) : 0.0f;
        g[i] = v;
    }
    ITEM_MAX(v)
}

// ratio = support ? data / max(est, eps) : 0 (prc.cu:270-274), in place. When the forward read
// the FP16 copy (`scaled`), est is hscale times the forward and is scaled back first (exactly:
// hscale is a power of two). pmax, when given, receives per-item maxima of the ratio.
__global__ void k_ratio(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                        const Scal* __restrict__ sc, const float* __restrict__ data, float* __restrict__ est,
                        double* __restrict__ pmax, int scaled) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const int lane = threadIdx.x & (kWave - 1);
    const Vol& V = vols[it.x];
    float r = 0.0f;
    if (lane < it.w) {
        const size_t loc = (size_t)it.y * V.nx + it.z + lane, i = V.vbase + loc;
        const float e = scaled ? est[i] / sc[it.x].hscale : est[i];
        r = gload(V.support + loc) ? data[i] / fmaxf(e, sc[it.x].eps) : 0.0f;
        est[i] = r;
    }
    ITEM_MAX(r)
}

// FP16 copy of a nonnegative vector times the volume's scale (which = 0: hscale, 1: rscale),
// rounded to nearest; saturation at the largest finite FP16 value only guards the bound.
__device__ inline __half to_half_sat(float v) { return __float2half_rn(fminf(v, 65504.0f)); }

__global__ void k_half_copy(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                            const Scal* __restrict__ sc, const float* __restrict__ src, __half* __restrict__ dst,
                            int which) {
    ITEM_PROLOGUE
    const float scale = which ? sc[it.x].rscale : sc[it.x].hscale;
    dst[i] = to_half_sat(src[i] * scale);
}

// RDP MAP-EM surrogate update (prc.cu:279-309), fused with the support mask and the FP64 partial
// sums of the count normalisation. One block of 256 threads updates a tile of 64 (x) x 4 (y)
// voxels over kTileZ slices, one voxel per thread per slice: it marches along z keeping three
// planes of f and the support, with a one-voxel halo, in an LDS ring (plane z + 1 is loaded
// while z - 1 .. z + 1 are read), so each value of f is read from memory 1.5 times instead of up
// to 27. The 26 distance weights min(s) / |(dx sx, dy sy, dz sz)| are computed once per block.
// Neighbours are visited in the reference's order (dz, dy, dx), and a neighbour outside the grid
// or the support (both stored as support 0 in the halo) is skipped, as prc.cu:290-293 does.
// Writes next on every voxel of the tile (0 off support) and one partial sum per tile.
constexpr int kTileY = 4;
constexpr int kTileZ = 4;
constexpr int kHaloX = kWave + 2, kHaloY = kTileY + 2, kHaloPlane = kHaloX * kHaloY;
static_assert(kWave * kTileY == 256, "one voxel per thread of a 256-thread block");

__global__ __launch_bounds__(256) void k_update(const Vol* __restrict__ vols, const int4* __restrict__ tiles,
                                                 const Scal* __restrict__ sc, const float* __restrict__ f,
                                                 const float* __restrict__ back, float* __restrict__ next,
                                                 double* __restrict__ partial) {
    __shared__ float fs[3][kHaloPlane];
    __shared__ unsigned char ss[3][kHaloPlane];
    __shared__ float wnb[27];
    __shared__ double red[256 / kWave];
    const int tid = threadIdx.x;
    const int4 t = tiles[blockIdx.x];
    const Vol& V = vols[t.x];
    const Scal& S = sc[t.x];
    const int nx = V.nx, ny = V.ny, nz = V.nz, x0 = t.y, y0 = t.z, z0 = t.w;
    const bool prior = S.beta != 0.0f;
    if (tid < 27) {
        const float smin = fminf(V.sx, fminf(V.sy, V.sz));
        const float mx = (tid % 3 - 1) * V.sx, my = (tid / 3 % 3 - 1) * V.sy, mz = (tid / 9 - 1) * V.sz;
        wnb[tid] = tid == 13 ? 0.0f : smin / sqrtf(mx * mx + my * my + mz * mz);
    }
    // Plane z of the tile with its halo into ring slot `slot` (outside the grid: support 0).
    auto load_plane = [&](int z, int slot) {
        for (int e = tid; e < kHaloPlane; e += 256) {
            const int x = x0 - 1 + e % kHaloX, y = y0 - 1 + e / kHaloX;
            const bool in = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
            const size_t loc = in ? ((size_t)z * ny + y) * nx + x : 0;
            const unsigned char su = in ? gload(V.support + loc) : (unsigned char)0;
            fs[slot][e] = su ? f[V.vbase + loc] : 0.0f;
            ss[slot][e] = su;
        }
    };
    const int lx = tid & (kWave - 1), ly = tid / kWave;
    const int x = x0 + lx, y = y0 + ly;
    const bool mine = x < nx && y < ny;
    const int c = (ly + 1) * kHaloX + lx + 1;  // the thread's voxel in a halo plane
    if (prior) {
        load_plane(z0 - 1, 0);
        load_plane(z0, 1);
    }
    double acc = 0.0;
    for (int k = 0; k < kTileZ && z0 + k < nz; ++k) {  // the bound is the block's own
        const int z = z0 + k;
        const int s_lo = k % 3, s_mid = (k + 1) % 3, s_hi = (k + 2) % 3;
        if (prior) load_plane(z + 1, s_hi);
        __syncthreads();  // the weights and this step's planes are in place
        float result = 0.0f;
        if (mine) {
            const size_t loc = ((size_t)z * ny + y) * nx + x, idx = V.vbase + loc;
            const bool su = prior ? ss[s_mid][c] != 0 : gload(V.support + loc) != 0;
            if (su) {
                const float fj = prior ? fs[s_mid][c] : f[idx];
                const float f_em = fj * (back[idx] * S.rinv);  // rinv: exact power of two
                if (!prior) {
                    result = fmaxf(f_em, 0.0f);
                } else {
                    const float gamma = S.gamma, eps = S.eps;
                    float grad = 0.0f, curv = 0.0f;
#pragma unroll
                    for (int dz = -1; dz <= 1; ++dz) {
                        const int slot = dz < 0 ? s_lo : dz > 0 ? s_hi : s_mid;
#pragma unroll
                        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                            for (int dx = -1; dx <= 1; ++dx) {
                                if (!(dx | dy | dz)) continue;
                                const int h = c + dy * kHaloX + dx;
                                if (!ss[slot][h]) continue;
                                const float w = wnb[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1];
                                const float fk = fs[slot][h], d = fj - fk, ad = fabsf(d);
                                // The hardware reciprocal (within 1 ulp; D >= eps > 0 is normal) in
                                // place of prc.cu's division: each term moves by a few 1e-8 relative.
                                const float D = fj + fk + gamma * ad + eps, invD = __builtin_amdgcn_rcpf(D);
                                grad += w * d * (fj + 3.0f * fk + gamma * ad + 2.0f * eps) * invD * invD;
                                const float half_quad = (2.0f * (fj + fk + eps) + gamma * ad) * invD * invD;
                                const float hess = 2.0f * (2.0f * fk + eps) * (2.0f * fk + eps) * invD * invD * invD;
                                curv += 2.0f * w * fmaxf(half_quad, hess);
                            }
                    }
                    const float a = S.beta * curv, B = 1.0f + S.beta * (grad - curv * fj);
                    const float root = sqrtf(B * B + 4.0f * a * f_em);
                    result = B >= 0.0f ? (f_em > 0.0f ? 2.0f * f_em / (B + root) : 0.0f) : (root - B) / (2.0f * a);
                }
            }
            next[idx] = result;
        }
        acc += (double)result;
        __syncthreads();  // every thread is done with slot s_lo before the next step refills it
    }
    acc = wave_sum_d(acc);
    if (lx == 0) red[ly] = acc;
    __syncthreads();
    if (tid == 0) {
        double tot = 0.0;
        for (int w = 0; w < 256 / kWave; ++w) tot += red[w];
        partial[blockIdx.x] = tot;
    }
}

// f = max(f, floor) on support (prc.cu:311-314) and per-item FP64 partial sums.
__global__ void k_initialize(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                             const Scal* __restrict__ sc, float* __restrict__ f, double* __restrict__ partial) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const int lane = threadIdx.x & (kWave - 1);
    const Vol& V = vols[it.x];
    float v = 0.0f;
    if (lane < it.w) {
        const size_t loc = (size_t)it.y * V.nx + it.z + lane, i = V.vbase + loc;
        v = f[i];
        if (gload(V.support + loc)) v = fmaxf(v, 1000.0f * sc[it.x].eps);
        f[i] = v;
    }
    const double tot = wave_sum_d((double)v);
    if (lane == 0) partial[item_id] = tot;
}

// next = RN(next * scale); with acceleration residual = next - residual (prc.cu:325-332), plus
// per-item partials of <h, h_prev> and <h_prev, h_prev>.
__global__ void k_finish(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                         const Scal* __restrict__ sc, float* __restrict__ next, float* __restrict__ residual,
                         const float* __restrict__ hprev, double* __restrict__ partial, int use_scale) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const int lane = threadIdx.x & (kWave - 1);
    const Vol& V = vols[it.x];
    double hy = 0.0, hh = 0.0;
    if (lane < it.w) {
        const size_t i = V.vbase + (size_t)it.y * V.nx + it.z + lane;
        const float value = next[i] * (use_scale ? sc[it.x].scale : 1.0f);
        next[i] = value;
        if (residual) {
            const float h = value - residual[i];
            residual[i] = h;
            const double hp = hprev[i];
            hy = (double)h * hp;
            hh = hp * hp;
        }
    }
    if (partial) {
        hy = wave_sum_d(hy);
        hh = wave_sum_d(hh);
        if (lane == 0) partial[2 * item_id] = hy, partial[2 * item_id + 1] = hh;
    }
}

// Per-volume FP64 totals of the item partials, fixed order (one block per volume).
// mode 0: initial count normalisation  scale = support_sum / sum
// mode 1: prior count normalisation    scale = sum > 0 ? support_sum / sum : 1 (beta > 0 only)
// mode 2: Biggs-Andrews lambda from (hy, hh) pairs
// mode 3: plain sum into Scal::sum
// mode 4, 5: largest item maximum -> hscale (4) or rscale and rinv (5), powers of two that put it
//            in [2^13, 2^14) (1 when it is 0)
__global__ void k_reduce_volume(const int* __restrict__ item_begin, const double* __restrict__ partial,
                                Scal* __restrict__ sc, int mode) {
    __shared__ double sh[2][256];
    const int v = blockIdx.x;
    const int b = item_begin[v], e = item_begin[v + 1];
    const int stride = mode == 2 ? 2 : 1;
    const bool is_max = mode >= 4;
    double a0 = 0.0, a1 = 0.0;
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) {
        const double p = partial[(size_t)stride * i];
        a0 = is_max ? fmax(a0, p) : a0 + p;
        if (mode == 2) a1 += partial[(size_t)stride * i + 1];
    }
    sh[0][threadIdx.x] = a0, sh[1][threadIdx.x] = a1;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] = is_max ? fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + o])
                                        : sh[0][threadIdx.x] + sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (is_max) {
        if (threadIdx.x) return;
        // m = mant * 2^ex with mant in [0.5, 1): m * 2^(14 - ex) lies in [2^13, 2^14).
        int ex = 0;
        const double m = sh[0][0];
        (void)frexp(m, &ex);
        const float scale = m > 0.0 ? ldexpf(1.0f, min(100, max(-100, 14 - ex))) : 1.0f;
        if (mode == 4) sc[v].hscale = scale;
        else sc[v].rscale = scale, sc[v].rinv = 1.0f / scale;
        return;
    }
    if (threadIdx.x) return;
    Scal& S = sc[v];
    const double s0 = sh[0][0], s1 = sh[1][0];
    if (mode == 0) {
        S.sum = s0;
        S.scale = (float)(S.support_sum / s0);
    } else if (mode == 1) {
        S.sum = s0;
        S.scale = (S.beta > 0.0f && s0 > 0.0) ? (float)(S.support_sum / s0) : 1.0f;
    } else if (mode == 2) {
        S.hy = s0, S.hh = s1;
        S.lambda = s1 > 0.0 ? (float)fmin(1.0, fmax(0.0, s0 / s1)) : 0.0f;
    } else {
        S.sum = s0;
    }
}

__global__ void k_sum_partial(const Vol* __restrict__ vols, const int4* __restrict__ items, int n_items,
                              const float* __restrict__ a, double* __restrict__ partial) {
    const int item_id = blockIdx.x * kItemsPerBlock + wave_in_block();
    if (item_id >= n_items) return;
    const int4 it = items[item_id];
    const int lane = threadIdx.x & (kWave - 1);
    const Vol& V = vols[it.x];
    double v = 0.0;
    if (lane < it.w) v = a[V.vbase + (size_t)it.y * V.nx + it.z + lane];
    v = wave_sum_d(v);
    if (lane == 0) partial[item_id] = v;
}

// ---------------------------------------------------------------------------------------
// Host side.

thread_local std::string g_error;

struct Failure {
    int code;
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    throw Failure{code};
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) fail(PRC_ERR_DEVICE, "HIP error '%s' at %s:%d", hipGetErrorString(e_), \
                                   __FILE__, __LINE__);                                           \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return PRC_OK;
    } catch (const Failure& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return PRC_ERR_MEMORY;
    }
}

template <class T>
T* dalloc(size_t n) {
    T* p = nullptr;
    if (n == 0) n = 1;
    const hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(PRC_ERR_MEMORY, "device allocation of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    }
    return p;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

inline unsigned blocks_for(long long n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------------------
// Optional per-launch timing with HIP events (prc_profile_enable). Graph replay is not used
// while it is on, so every kernel launch is bracketed on its own stream.
struct ProfRecord {
    const char* name;
    hipEvent_t begin, end;
};
bool g_prof_on = false;
std::vector<ProfRecord> g_prof;

struct ProfScope {
    ProfRecord rec{};
    hipStream_t stream;
    bool on;
    ProfScope(const char* name, hipStream_t s) : stream(s), on(g_prof_on) {
        if (!on) return;
        rec.name = name;
        HIPCHK(hipEventCreate(&rec.begin));
        HIPCHK(hipEventCreate(&rec.end));
        HIPCHK(hipEventRecord(rec.begin, stream));
    }
    ~ProfScope() {
        if (!on) return;
        if (hipEventRecord(rec.end, stream) == hipSuccess) g_prof.push_back(rec);
    }
};

}  // namespace

// One volume's projector.
struct prc_projector {
    prc_geometry g{};
    Isotope model{};
    Vol hv{};
    Vol* dv = nullptr;
    float* rho_alloc = nullptr;  // padded density with its kHalo margin
    float* rho_p = nullptr;      // first voxel of the padded volume inside rho_alloc
    float* table = nullptr;
    unsigned char* support = nullptr;
    float* W = nullptr;
    int* row_xa = nullptr;
    int* row_len = nullptr;
    RowInfo* rinfo = nullptr;
    RowInfo* tinfo = nullptr;
    void* wraw = nullptr;  // allocations of the two weight copies (with slack)
    void* traw = nullptr;
    void* wbuf = nullptr;  // source-major planes = wraw + one wave
    void* tbuf = nullptr;  // target-major planes = traw + one wave
    long long wcap = 0, tcap = 0;  // weight buffer capacities (entries)
    long long margin = -1;         // elements of padding on each side of gathered vectors
    FarStep* fstep = nullptr;
    unsigned* fsoa = nullptr;
    int4* pstep = nullptr;
    int* plin = nullptr;
    int* kfirst = nullptr;
    NearStep* nstep = nullptr;
    int* nlen = nullptr;
    int* kglob = nullptr;
    long long* row_tot = nullptr;
    long long* row_base = nullptr;
    long long* unit_planes = nullptr;  // per segment: list length (aligned)
    long long* tile_sums = nullptr;    // layout scan: sums of segment tiles (two arrays)
    int* cnt_t = nullptr;              // [n_rays + 1][segments] source-major counts, ray-major
    int* kt_t = nullptr;               // [n_rays + 1][segments] forward bounds kt(t, p), ray-major
    TargetRun* tmaj = nullptr;         // [n_rays + 1][segments] target-major runs, ray-major
    long long* list_base = nullptr;
    long long* total = nullptr;  // entries (source, target), list lengths (source, target)
    PlaneRun* sruns = nullptr;  // per segment: plane sequence of each copy
    PlaneRun* truns = nullptr;
    int* slist = nullptr;  // shift lists of the two copies (with kListSlack zero entries)
    int* tlist = nullptr;
    long long slcap = 0, tlcap = 0;  // list capacities (entries)
    int4* items_act = nullptr;  // active row segments (march, adjoint)
    int4* items_all = nullptr;  // every row in 64-voxel pieces (streaming kernels)
    int4* items_fwd = nullptr;  // active segments plus the zero pieces around them (forward)
    int4* tiles = nullptr;      // 64 x 4 x kTileZ voxel tiles of the RDP update (vol, x0, y0, z0)
    int n_tiles = 0;
    int* queue = nullptr;       // work-queue counters of the projector passes
    // A normalisation runs on the library's build stream, or on the stream of the batch that asks
    // for it (prc_batch_normalise). It first waits for the null stream (solvers) and for the
    // streams of the other batches that use this projector.
    std::vector<hipStream_t> users;
    // Pinned staging of the sizes normalise_begin reads back (so the read-back does not block
    // the host): total[4] (long long), then the memo counts [2] and kglob [n_rays + 1] (int).
    void* h_stage = nullptr;
    PassShape fwd_shape = kShapes[kDefaultShape], adj_shape = kShapes[kDefaultShape];  // tuned with the first solver
    bool counted = false;  // march statistics read (after the first normalisation)
    bool tcleared = false;  // the count pass of the current normalisation zeroed traw (all of tcap)
    bool tuned = false;
    // Far rays re-marched in every pass instead of stored: set by PRC_FLAG_MARCH_PASSES, or by the
    // first normalisation when both stored copies would not fit in device memory.
    bool far_marched = false;
    int n_act = 0, n_all = 0, n_fwd = 0;
    std::vector<int4> h_items_act, h_items_all, h_items_fwd, h_tiles;
    std::vector<int> h_row_xa, h_row_len;  // active x range of every row
    std::vector<unsigned char> h_support;
    unsigned long long* counters = nullptr;  // vrs_near, vrs_far, deposits
    // Near-block memoisation (uniform 5x5x5 neighbourhoods): per-voxel class (-2 inactive,
    // -1 marched, t >= 0 template t), the marched list, and the templates.
    bool memo = false;
    signed char* cls = nullptr;
    int2* near_list = nullptr;
    int n_list = 0;
    float* tmpl = nullptr;
    float* tmpl_sum = nullptr;
    int* memo_meta = nullptr;  // template keys [64], representatives [64], list and uniform counts
    float* scratch_in = nullptr;   // margin-padded input of the device-pointer forward
    float* scratch_adj = nullptr;  // margin-padded input of the device-pointer adjoint
    prc_stats stats{};
    size_t n = 0;
    int wbytes = 4;
};

struct prc_solver {
    prc_projector* A = nullptr;
    prc_solve_params p{};
    float *g = nullptr, *f = nullptr, *prev = nullptr, *y = nullptr, *hprev = nullptr, *est = nullptr,
          *gf = nullptr;
    float* base[5] = {};  // f, prev, y, hprev, est as allocated (the loop rotates them)
    float* raw[7] = {};   // allocations: every vector has the projector's gather margin on both sides
    __half* hraw[2] = {};  // mixed form: FP16 forward and adjoint inputs, same layout
    Scal* sc = nullptr;
    Scal sc0{};
    double* partial = nullptr;
    int* item_begin = nullptr;
    int* tile_begin = nullptr;
    double support_sum = 0.0;
    bool initialised = false;
};

namespace {

// Row-segment work items: 64 consecutive voxels of one X row. Lists are padded to a multiple of
// kItemsPerBlock with empty items so that a block never mixes volumes.
void build_items(prc_projector* A) {
    const int nx = A->g.nx, rows = A->g.ny * A->g.nz;
    A->h_items_act.clear();
    A->h_items_all.clear();
    A->h_items_fwd.clear();
    auto zeros = [&](int row, int b, int e) {  // [b, e) of a row, as zero pieces of the forward
        for (int x0 = b; x0 < e; x0 += kWave) A->h_items_fwd.push_back(make_int4(0, row, x0, -std::min(kWave, e - x0)));
    };
    std::vector<int> xa(rows, 0), len(rows, 0);
    for (int row = 0; row < rows; ++row) {
        const unsigned char* s = A->h_support.data() + (size_t)row * nx;
        int lo = -1, hi = -1;
        for (int x = 0; x < nx; ++x)
            if (s[x]) {
                if (lo < 0) lo = x;
                hi = x;
            }
        if (lo >= 0) {
            xa[row] = lo, len[row] = hi - lo + 1;
            zeros(row, 0, lo);
            for (int x0 = lo; x0 <= hi; x0 += kWave) {
                A->h_items_act.push_back(make_int4(0, row, x0, std::min(kWave, hi + 1 - x0)));
                A->h_items_fwd.push_back(A->h_items_act.back());
            }
            zeros(row, hi + 1, nx);
        } else {
            zeros(row, 0, nx);
        }
        for (int x0 = 0; x0 < nx; x0 += kWave) A->h_items_all.push_back(make_int4(0, row, x0, std::min(kWave, nx - x0)));
    }
    while (A->h_items_act.size() % kItemsPerBlock) A->h_items_act.push_back(make_int4(0, 0, 0, 0));
    while (A->h_items_all.size() % kItemsPerBlock) A->h_items_all.push_back(make_int4(0, 0, 0, 0));
    while (A->h_items_fwd.size() % kItemsPerBlock) A->h_items_fwd.push_back(make_int4(0, 0, 0, 0));
    A->n_act = (int)A->h_items_act.size();
    A->n_all = (int)A->h_items_all.size();
    A->n_fwd = (int)A->h_items_fwd.size();
    A->h_row_xa = xa, A->h_row_len = len;
    A->row_xa = dalloc<int>(rows);
    A->row_len = dalloc<int>(rows);
    HIPCHK(hipMemcpy(A->row_xa, xa.data(), rows * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(A->row_len, len.data(), rows * sizeof(int), hipMemcpyHostToDevice));
    A->items_act = dalloc<int4>(A->n_act);
    A->items_all = dalloc<int4>(A->n_all);
    HIPCHK(hipMemcpy(A->items_act, A->h_items_act.data(), A->n_act * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(A->items_all, A->h_items_all.data(), A->n_all * sizeof(int4), hipMemcpyHostToDevice));
    A->items_fwd = dalloc<int4>(A->n_fwd);
    A->queue = dalloc<int>(kXcds);
    HIPCHK(hipMemcpy(A->items_fwd, A->h_items_fwd.data(), A->n_fwd * sizeof(int4), hipMemcpyHostToDevice));
    long long n_rows = 0;
    for (int row = 0; row < rows; ++row) n_rows += len[row] > 0;
    A->stats.n_rows = n_rows;
    A->h_tiles.clear();
    for (int z = 0; z < A->g.nz; z += kTileZ)
        for (int y = 0; y < A->g.ny; y += kTileY)
            for (int x = 0; x < nx; x += kWave) A->h_tiles.push_back(make_int4(0, x, y, z));
    A->n_tiles = (int)A->h_tiles.size();
    A->tiles = dalloc<int4>(A->n_tiles);
    HIPCHK(hipMemcpy(A->tiles, A->h_tiles.data(), A->n_tiles * sizeof(int4), hipMemcpyHostToDevice));
}

template <class WT>
void launch_far(prc_projector* A, int pass, hipStream_t st) {
    const bool exact = A->hv.exact, count = !A->counted;
    const unsigned nb = blocks_for(A->n_act, kItemsPerBlock);
    if (pass == 1) {
        ProfScope ps("far_count", st);
        // A rebuild zeroes the target-major copy it will refill during the count (the fill writes
        // only the entries that receive a deposit); normalise_finish clears it itself when the
        // buffer is new or grows.
        A->tcleared = A->traw != nullptr && !A->far_marched;
        v4i* clear = A->tcleared ? (v4i*)A->traw : nullptr;
        const long long units = A->tcleared ? (A->tcap + 2 * kWave) * A->wbytes / 16 : 0;
        if (count)
            k_far_count<true><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, A->fstep, A->kfirst, A->rinfo,
                                                  A->counters + 1, clear, units);
        else
            k_far_count<false><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, A->fstep, A->kfirst, A->rinfo,
                                                   nullptr, clear, units);
    } else {
        ProfScope ps("far_fill", st);
        unsigned long long* ndep = A->counters + 2;
        if (exact && count) k_far_fill<true, WT, true><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, ndep);
        else if (exact) k_far_fill<true, WT, false><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, nullptr);
        else if (count) k_far_fill<false, WT, true><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, ndep);
        else k_far_fill<false, WT, false><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, nullptr);
    }
    HIPCHK(hipGetLastError());
}

template <class WT>
void launch_near(prc_projector* A, hipStream_t st) {
    const unsigned nb = blocks_for(A->n_act, kItemsPerBlock);
    ProfScope ps("near", st);
    // The voxel visits are counted by a projector's first build only.
    const bool count = !A->counted;
    unsigned long long* vrs = A->counters;
    auto near_list = [&](auto exact_c, auto count_c) {
        constexpr bool E = decltype(exact_c)::value, C = decltype(count_c)::value;
        k_near_list<E, WT, C><<<blocks_for(A->n_list, 256), 256, 0, st>>>(A->dv, A->near_list, A->n_list,
                                                                                A->tmpl, A->tmpl_sum, vrs);
    };
    auto near_rows = [&](auto exact_c, auto count_c) {
        constexpr bool E = decltype(exact_c)::value, C = decltype(count_c)::value;
        k_near<E, WT, C><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, vrs);
    };
    using T = std::true_type;
    using F = std::false_type;
    if (A->memo) {  // march the listed sources, then copy the templates to the uniform ones
        if (A->hv.exact) count ? near_list(T(), T()) : near_list(T(), F());
        else count ? near_list(F(), T()) : near_list(F(), F());
        k_near_fill<WT><<<nb, 256, 0, st>>>(A->dv, A->items_act, A->n_act, A->cls, A->tmpl, A->tmpl_sum);
    } else if (A->hv.exact) {
        count ? near_rows(T(), T()) : near_rows(T(), F());
    } else {
        count ? near_rows(F(), T()) : near_rows(F(), F());
    }
    HIPCHK(hipGetLastError());
}

// A far-march pass (k_far_march) over the projector's active segments, described in table dv (the
// projector's own, or a batch's, whose items then carry the batch index of the volume).
template <class WT, class VT>
void launch_far_march(const Vol* dv, const prc_projector* A, int mode, const int4* items, const void* in, float* out,
                      hipStream_t st, int n_items = -1) {
    if (!items) items = A->items_act, n_items = A->n_act;
    const unsigned nb = blocks_for(n_items, kItemsPerBlock);
    const VT* v = (const VT*)in;
    ProfScope ps(mode == kMarchNorm ? "far_fill" : mode == kMarchAdj ? "adjoint_far_march" : "forward_far_march", st);
    auto go = [&](auto exact_c) {
        constexpr bool E = decltype(exact_c)::value;
        if (mode == kMarchNorm) k_far_march<E, WT, VT, kMarchNorm><<<nb, 256, 0, st>>>(dv, items, n_items, v, out);
        else if (mode == kMarchAdj) k_far_march<E, WT, VT, kMarchAdj><<<nb, 256, 0, st>>>(dv, items, n_items, v, out);
        else k_far_march<E, WT, VT, kMarchFwd><<<nb, 256, 0, st>>>(dv, items, n_items, v, out);
    };
    if (A->hv.exact) go(std::true_type());
    else go(std::false_type());
    HIPCHK(hipGetLastError());
}

// Density upload, support, work items, range table and ray tables (everything that does not
// depend on the march).
void setup_projector(prc_projector* A, const float* host_rho, const float* host_table) {
    const prc_geometry& g = A->g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    A->n = n;
    A->model = renormalised(g.isotope);
    const float cutoff = range_cutoff_host(A->model, 0.9995f);  // prc.cu:587
    A->stats.cutoff_mm = cutoff;
    A->stats.n_voxels = (long long)n;
    A->wbytes = g.precision == PRC_PRECISION_MIXED ? 2 : 4;

    // Density: clamp (prc.cu:603), support mask (prc.cu:605), kHalo halo of width one voxel.
    const int nxp = g.nx + 2, nyp = g.ny + 2, nzp = g.nz + 2;
    const size_t np = (size_t)nxp * nyp * nzp;
    // kHalo margin on both sides for the march's group loads: every lane of a wave loads the
    // group's voxels, and a group runs only while some lane of the wave is alive. That lane's
    // last voxel is inside the grid and a group moves it at most kMarchGroup voxels, at most
    // kMarchGroup padded slices; the other lanes of the wave are at most 63 voxels further along
    // the row. So (kMarchGroup + 1) slices and two waves of margin hold every group load.
    const size_t margin = (size_t)(kMarchGroup + 1) * nxp * nyp + 2 * kWave;
    if ((np + 2 * margin) * sizeof(float) > (size_t)INT_MAX)
        fail(PRC_ERR_INVALID, "Volume too large for 32-bit density addressing");
    std::vector<float> rp(np + 2 * margin, kHalo);
    A->h_support.assign(n, 0);
    long long n_active = 0;
    int box[6] = {g.nx, -1, g.ny, -1, g.nz, -1};
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j) {
            const size_t src = ((size_t)k * g.ny + j) * g.nx;
            float* dst = rp.data() + margin + (((size_t)(k + 1) * nyp + (j + 1)) * nxp + 1);
            for (int i = 0; i < g.nx; ++i) {
                const float r = std::max(host_rho[src + i], kMinDensity);
                dst[i] = r;
                const bool s = r >= g.support_rho;
                A->h_support[src + i] = s;
                n_active += s;
                if (s) {
                    box[0] = std::min(box[0], i), box[1] = std::max(box[1], i);
                    box[2] = std::min(box[2], j), box[3] = std::max(box[3], j);
                    box[4] = std::min(box[4], k), box[5] = std::max(box[5], k);
                }
            }
        }
    A->stats.n_active = n_active;
    A->far_marched = (g.flags & PRC_FLAG_MARCH_PASSES) != 0;
    if (!(g.flags & PRC_FLAG_NO_NEAR_MEMO)) {
        A->cls = dalloc<signed char>(n);
        A->near_list = dalloc<int2>(std::max<long long>(n_active, 1));
        A->tmpl = dalloc<float>((size_t)kMaxTemplates * kNQ);
        A->tmpl_sum = dalloc<float>(kMaxTemplates);
        A->memo_meta = dalloc<int>(2 * kMaxTemplates + 2);
    }
    A->rho_alloc = dalloc<float>(np + 2 * margin);
    HIPCHK(hipMemcpy(A->rho_alloc, rp.data(), (np + 2 * margin) * sizeof(float), hipMemcpyHostToDevice));
    A->rho_p = A->rho_alloc + margin;
    std::vector<float>().swap(rp);
    A->support = dalloc<unsigned char>(n);
    HIPCHK(hipMemcpy(A->support, A->h_support.data(), n, hipMemcpyHostToDevice));
    A->W = dalloc<float>(n);
    A->counters = dalloc<unsigned long long>(4);

    // Range table on the device (prc.cu:625-633) unless the caller supplies one.
    const float scale = (float)kBins / cutoff;
    A->table = dalloc<float>(kBins + 1);
    if (host_table) {
        HIPCHK(hipMemcpy(A->table, host_table, (kBins + 1) * sizeof(float), hipMemcpyHostToDevice));
    } else {
        Branch b[3] = {};
        for (int i = 0; i < A->model.n; ++i) b[i] = A->model.br[i];
        k_range_table<<<blocks_for(kBins + 1, 256), 256>>>(A->table, b[0], b[1], b[2], A->model.n, scale);
        HIPCHK(hipGetLastError());
    }

    // Ray and step tables.
    const int R = g.n_rays, R1 = R + 1;
    const int L = g.nx + g.ny + g.nz;  // a ray leaves any grid within nx+ny+nz voxel visits
    const int LP = std::max(L, kNQ);
    const int LF = (int)align_up(L + kFarGroup, kMarchGroup);  // the fill reads up to kFarGroup - 1 past L
    A->fstep = dalloc<FarStep>((size_t)R * LF);
    A->fsoa = dalloc<unsigned>((size_t)R * 2 * LF);
    A->pstep = dalloc<int4>((size_t)R1 * LP);
    A->plin = dalloc<int>((size_t)R1 * LP);
    HIPCHK(hipMemset(A->plin, 0, (size_t)R1 * LP * sizeof(int)));
    A->kfirst = dalloc<int>(R1);
    A->kglob = dalloc<int>(R1);
    A->nstep = dalloc<NearStep>((size_t)kNearRays * kNearMax);
    A->nlen = dalloc<int>(kNearRays);
    HIPCHK(hipMemset(A->nstep, 0, (size_t)kNearRays * kNearMax * sizeof(NearStep)));

    Vol& V = A->hv;
    V.nx = g.nx, V.ny = g.ny, V.nz = g.nz, V.n_rays = R;
    V.sx = g.sx, V.sy = g.sy, V.sz = g.sz, V.cutoff = cutoff, V.support_rho = g.support_rho;
    V.air_cap = g.air_cap_mm, V.scale = scale;
    V.nb = A->model.n;
    for (int i = 0; i < 3; ++i) V.br[i] = i < A->model.n ? A->model.br[i] : Branch{0, 0, 0, 0, 0};
    V.exact = (g.flags & PRC_FLAG_EXACT_RANGE) != 0;
    V.L = L, V.LP = LP, V.LF = LF, V.vbase = 0;
    std::copy(box, box + 6, V.box);
    V.rho_p = A->rho_p, V.table = A->table, V.support = A->support, V.W = A->W;
    V.rho_base = A->rho_alloc, V.rho_bytes = (long long)(np + 2 * margin) * (long long)sizeof(float);
    V.fstep = A->fstep, V.fsoa = A->fsoa, V.pstep = A->pstep, V.plin = A->plin, V.kfirst = A->kfirst, V.nstep = A->nstep;
    V.nlen = A->nlen;

    k_rays<<<blocks_for(R + kNearRays + kNQ, 128), 128>>>(V, A->fstep, A->fsoa, A->pstep, A->plin, A->kfirst,
                                                          A->nstep, A->nlen);
    HIPCHK(hipGetLastError());

    build_items(A);
    V.row_xa = A->row_xa, V.row_len = A->row_len;
    V.nseg = (g.nx + kWave - 1) / kWave;
    const size_t units = (size_t)g.ny * g.nz * V.nseg;
    A->rinfo = dalloc<RowInfo>(units * R1);
    A->tinfo = dalloc<RowInfo>(units * R1);
    A->row_tot = dalloc<long long>(units);
    A->row_base = dalloc<long long>(units);
    A->unit_planes = dalloc<long long>(units);
    A->tile_sums = dalloc<long long>(2 * blocks_for((long long)units, kScanTile));
    A->cnt_t = dalloc<int>(units * R1);
    A->kt_t = dalloc<int>(units * R1);
    A->tmaj = dalloc<TargetRun>(units * R1);
    A->list_base = dalloc<long long>(units);
    A->sruns = dalloc<PlaneRun>(units);
    A->truns = dalloc<PlaneRun>(units);
    A->total = dalloc<long long>(4);
    V.rinfo = A->rinfo, V.tinfo = A->tinfo;
    V.sruns = A->sruns, V.truns = A->truns, V.tmaj = A->tmaj;
    V.wbuf = nullptr, V.tbuf = nullptr, V.slist = nullptr, V.tlist = nullptr;
    A->dv = dalloc<Vol>(1);
    HIPCHK(hipMemcpy(A->dv, &V, sizeof(Vol), hipMemcpyHostToDevice));
}

// Grow a weight buffer to hold `entries`, with one wave of slack on both sides: the operators
// read whole 64-lane segments, and a forward segment may start up to 63 voxels before its row's
// active range (x0 - xa >= -63) or end up to 63 past the last plane. `raw` is the allocation,
// `buf` = raw + kWave elements.
void ensure_weights(void*& raw, void*& buf, long long& cap, long long entries, int wbytes) {
    const long long need = entries;
    if (raw && need <= cap) return;
    if (raw) (void)hipFree(raw);
    raw = buf = nullptr, cap = 0;
    raw = (void*)dalloc<unsigned char>((size_t)(need + 2 * kWave) * wbytes);
    buf = (unsigned char*)raw + (size_t)kWave * wbytes;
    cap = need;
}

// row_base / list_base = exclusive prefix sums of row_tot / unit_planes over the segments.
void scan_units(prc_projector* A, long long* total_a, long long* total_b, hipStream_t st) {
    const int units = A->hv.ny * A->hv.nz * A->hv.nseg;
    const int tiles = (int)blocks_for(units, kScanTile);
    long long* ta = A->tile_sums;
    long long* tb = A->tile_sums + tiles;
    k_tile_sums<<<tiles, kScanTile, 0, st>>>(A->row_tot, A->unit_planes, units, ta, tb);
    k_scan_tiles<<<1, kScanTile, 0, st>>>(ta, tb, tiles, total_a, total_b);
    k_scan_in_tiles<<<tiles, kScanTile, 0, st>>>(A->row_tot, A->unit_planes, units, ta, tb, A->row_base,
                                                        A->list_base);
}

// Shift lists, with kListSlack entries of slack past the last list. Allocated zeroed; entries past
// the current lists hold zeros or shifts of an earlier normalisation, valid shifts either way (the
// gather margin never changes after the first normalisation).
void ensure_list(int*& list, long long& cap, long long entries) {
    if (list && entries <= cap) return;
    if (list) (void)hipFree(list);
    list = nullptr, cap = 0;
    list = dalloc<int>((size_t)(entries + kListSlack));
    HIPCHK(hipMemset(list, 0, (size_t)(entries + kListSlack) * sizeof(int)));
    cap = entries;
}

// The normalisation pass (Projector::Projector, prc.cu:405-422): count, lay out both copies of
// the matrix, then march again writing every weight (source-major for the adjoint, target-major
// for the forward) and W. One host synchronisation (the buffer sizes); buffers are reused.
// Phase 1, enqueued on the projector's stream: counts, layout, and the read-back of the sizes into
// the projector's pinned staging (normalise_finish reads them after a stream sync).
void layout_and_sizes(prc_projector* A, hipStream_t st);

void normalise_begin(prc_projector* A, hipStream_t st) {
    Vol& V = A->hv;
    const size_t units = (size_t)V.ny * V.nz * V.nseg;
    const int R1 = V.n_rays + 1;
    HIPCHK(hipStreamSynchronize(nullptr));  // solver work on this projector
    for (hipStream_t u : A->users)          // other batches on this projector (st itself is in order)
        if (u != st) HIPCHK(hipStreamSynchronize(u));
    HIPCHK(hipMemsetAsync(A->rinfo, 0, units * R1 * sizeof(RowInfo), st));
    HIPCHK(hipMemsetAsync(A->counters, 0, 4 * sizeof(unsigned long long), st));
    HIPCHK(hipMemsetAsync(A->W, 0, A->n * sizeof(float), st));
    if (A->cls) {
        ProfScope ps("near_classify", st);
        int* keys = A->memo_meta;
        HIPCHK(hipMemsetAsync(keys, 0xff, kMaxTemplates * sizeof(int), st));
        HIPCHK(hipMemsetAsync(keys + kMaxTemplates, 0x7f, kMaxTemplates * sizeof(int), st));
        HIPCHK(hipMemsetAsync(keys + 2 * kMaxTemplates, 0, 2 * sizeof(int), st));
        const unsigned nb = blocks_for((long long)A->n, 256);
        k_classify<<<nb, 256, 0, st>>>(A->dv, A->cls, (unsigned*)keys, keys + kMaxTemplates);
        k_memo_list<<<blocks_for((long long)A->n, kMemoChunks * kMemoBlock), kMemoBlock, 0, st>>>(
            (int)A->n, A->cls, keys + kMaxTemplates, A->near_list, keys + 2 * kMaxTemplates);
        HIPCHK(hipGetLastError());
    }
    // The count pass sizes the stored layout; a projector whose far rays are marched in its passes
    // runs it only once, for the statistics, and stores the near block alone.
    if (!A->far_marched || !A->counted) {
        if (A->wbytes == 2) launch_far<__half>(A, 1, st);
        else launch_far<float>(A, 1, st);
    }
    layout_and_sizes(A, st);
}

// Both layouts and the read-back of their sizes into the projector's pinned staging (from the
// counts in rinfo; with far_marched, the far counts are cleared first).
void layout_and_sizes(prc_projector* A, hipStream_t st) {
    Vol& V = A->hv;
    const size_t units = (size_t)V.ny * V.nz * V.nseg;
    const int R1 = V.n_rays + 1;
    if (A->far_marched) {
        k_clear_far_counts<<<blocks_for((long long)units * R1, 256), 256, 0, st>>>(A->rinfo, (long long)units, R1);
        HIPCHK(hipGetLastError());
    }
    {
        ProfScope ps("layout", st);
        const unsigned nr = blocks_for((long long)units * R1, 256);
        const unsigned nw = blocks_for((long long)units, 256 / kWave);  // one wave per segment
        k_row_totals<<<nw, 256, 0, st>>>(V, A->row_tot, A->unit_planes, A->sruns, A->cnt_t);
        scan_units(A, A->total, A->total + 2, st);
        k_row_offsets<<<nr, 256, 0, st>>>(V, A->row_base, A->list_base, A->sruns);
        HIPCHK(hipMemsetAsync(A->kglob, 0, R1 * sizeof(int), st));
        k_kglob<<<R1 * kKglobChunks, 256, 0, st>>>(V, A->cnt_t, A->kglob);
        k_forward_bounds<<<nr, 256, 0, st>>>(V, A->kglob, A->cnt_t, A->kt_t);
        k_target_totals<<<nw, 256, 0, st>>>(V, A->row_tot, A->unit_planes, A->truns, A->kt_t);
        scan_units(A, A->total + 1, A->total + 3, st);
        k_target_offsets<<<nr, 256, 0, st>>>(V, A->row_base, A->list_base, A->truns, A->tmaj);
        HIPCHK(hipGetLastError());
    }
    if (!A->h_stage) HIPCHK(hipHostMalloc(&A->h_stage, 4 * sizeof(long long) + (2 + R1) * sizeof(int)));
    long long* total = (long long*)A->h_stage;
    int* ctr = (int*)(total + 4);
    HIPCHK(hipMemcpyAsync(total, A->total, 4 * sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ctr + 2, A->kglob, R1 * sizeof(int), hipMemcpyDeviceToHost, st));
    if (A->cls) HIPCHK(hipMemcpyAsync(ctr, A->memo_meta + 2 * kMaxTemplates, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
}

// Phase 2, after phase 1 has completed: buffers, then the near block and the fill, enqueued on the
// projector's stream.
// Device bytes of both stored copies and their shift lists for the sizes in `total`.
long long stored_bytes(const long long* total, int wbytes) {
    return (total[0] + total[1] + 4 * kWave) * wbytes + (total[2] + total[3] + 2 * kListSlack) * (long long)sizeof(int);
}

// Whether the stored matrix of these sizes fits: yes when the projector's buffers already hold
// them; else PRC_WEIGHT_BUDGET (bytes) when set, else the free device memory plus what the
// projector already holds, less a reserve for the solver's vectors and other allocations (2 GiB or
// 5 % of the device, whichever is larger).
bool stored_matrix_fits(const prc_projector* A, const long long* total) {
    // A rebuild into buffers that already hold these sizes needs no memory at all.
    if (A->wraw && A->traw && A->slist && A->tlist && total[0] <= A->wcap && total[1] <= A->tcap &&
        total[2] <= A->slcap && total[3] <= A->tlcap)
        return true;
    const long long need = stored_bytes(total, A->wbytes);
    if (const char* env = getenv("PRC_WEIGHT_BUDGET")) return need <= atoll(env);
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    long long held = 0;
    if (A->wraw) held += (A->wcap + 2 * kWave) * A->wbytes;
    if (A->traw) held += (A->tcap + 2 * kWave) * A->wbytes;
    if (A->slist) held += (A->slcap + kListSlack) * (long long)sizeof(int);
    if (A->tlist) held += (A->tlcap + kListSlack) * (long long)sizeof(int);
    const long long reserve = std::max<long long>(2LL << 30, (long long)(total_b / 20));
    return need <= (long long)free_b + held - reserve;
}

void normalise_finish(prc_projector* A, hipStream_t st) {
    Vol& V = A->hv;
    const int R1 = V.n_rays + 1;
    const long long* total = (const long long*)A->h_stage;
    if (!A->far_marched && !stored_matrix_fits(A, total)) {
        // The first build that does not fit switches the projector to marching its far rays in
        // every pass; the near block alone is laid out and stored.
        if (A->margin >= 0) fail(PRC_ERR_MEMORY, "the stored matrix no longer fits in device memory");
        A->far_marched = true;
        layout_and_sizes(A, st);
        HIPCHK(hipStreamSynchronize(st));
    }
    A->stats.far_marched = A->far_marched ? 1 : 0;
    const int* ctr = (const int*)(total + 4);
    const std::vector<int> kglob(ctr + 2, ctr + 2 + R1);
    if (A->cls) {
        A->n_list = ctr[0];
        A->stats.near_uniform = ctr[1];
        // Worth a second pass only when a good share of the sources is uniform.
        A->memo = ctr[1] > 0 && (long long)ctr[1] * 10 >= A->stats.n_active;
        if (!A->memo) A->stats.near_uniform = 0;
    }
    // Every stored step moves at most maxk voxels along each axis from its source.
    const int maxk = *std::max_element(kglob.begin(), kglob.end());
    const long long margin = (long long)(maxk + 1) * ((long long)V.nx * V.ny + V.nx + 1) + kWave;
    if (A->margin < 0) A->margin = margin;
    else if (margin > A->margin) fail(PRC_ERR_INVALID, "the system matrix reach grew after the first normalisation");
    ensure_weights(A->wraw, A->wbuf, A->wcap, total[0], A->wbytes);
    const long long tcap_before = A->tcap;  // ensure_weights changes the capacity when it reallocates
    ensure_weights(A->traw, A->tbuf, A->tcap, total[1], A->wbytes);
    const bool tzero = A->tcleared && A->tcap == tcap_before;  // zeroed by the count pass, buffer kept
    A->tcleared = false;
    if (total[2] > INT_MAX - kListSlack || total[3] > INT_MAX - kListSlack)
        fail(PRC_ERR_INVALID, "the system matrix holds too many planes for 32-bit shift-list offsets");
    ensure_list(A->slist, A->slcap, total[2]);
    ensure_list(A->tlist, A->tlcap, total[3]);
    A->stats.weight_entries = total[0];
    A->stats.target_entries = total[1];
    A->stats.weight_bytes = (total[0] + total[1]) * A->wbytes;
    A->stats.weight_planes = total[2];
    A->stats.target_planes = total[3];
    V.wbuf = A->wbuf, V.tbuf = A->tbuf, V.slist = A->slist, V.tlist = A->tlist;
    HIPCHK(hipMemcpyAsync(A->dv, &V, sizeof(Vol), hipMemcpyHostToDevice, st));
    // The fill writes only the target-major entries that receive a deposit: the rest must be zero.
    if (!tzero) HIPCHK(hipMemsetAsync(A->traw, 0, (size_t)(A->tcap + 2 * kWave) * A->wbytes, st));
    {
        const size_t units = (size_t)V.ny * V.nz * V.nseg;
        const unsigned nb = blocks_for((long long)units, 256 / kWave);
        const int pairs = A->wbytes == 2;
        k_plane_lists<<<nb, 256, 0, st>>>(V, A->rinfo, A->sruns, 1, pairs, A->slist);
        k_plane_lists<<<nb, 256, 0, st>>>(V, A->tinfo, A->truns, -1, pairs, A->tlist);
        HIPCHK(hipGetLastError());
    }
    if (A->wbytes == 2) {
        launch_near<__half>(A, st);
        A->far_marched ? launch_far_march<__half, __half>(A->dv, A, kMarchNorm, nullptr, nullptr, nullptr, st)
                       : launch_far<__half>(A, 2, st);
    } else {
        launch_near<float>(A, st);
        A->far_marched ? launch_far_march<float, float>(A->dv, A, kMarchNorm, nullptr, nullptr, nullptr, st)
                       : launch_far<float>(A, 2, st);
    }
    if (!A->counted) {  // statistics of the first build: the stored weights that are not zero
        const long long words = total[0] * A->wbytes / 4;
        if (A->wbytes == 2) k_count_nonzero<__half><<<2048, 256, 0, st>>>((const unsigned*)A->wbuf, words, A->counters + 3);
        else k_count_nonzero<float><<<2048, 256, 0, st>>>((const unsigned*)A->wbuf, words, A->counters + 3);
        HIPCHK(hipGetLastError());
    }
}

// The library's build stream: normalisations that no batch asks for run here, one after another
// (each march kernel fills the device by itself). Created on first use; the library keeps few
// streams so that a batch's stream gets a hardware queue of its own and its work overlaps the
// builds instead of queueing behind them.
hipStream_t g_build_stream = nullptr;
hipStream_t build_stream() {
    if (!g_build_stream) HIPCHK(hipStreamCreateWithFlags(&g_build_stream, hipStreamNonBlocking));
    return g_build_stream;
}

// The normalisation pass (Projector::Projector, prc.cu:405-422) of several projectors, enqueued on
// stream st: every projector's count and layout, one host synchronisation for the buffer sizes,
// then every projector's near block and fill. With `wait` the call returns when the builds are
// done; otherwise they stay queued on st (a batch's iterations queued after them see the result).
void read_counters(prc_projector* A);

void normalise_projectors(prc_projector* const* As, int n, hipStream_t st, bool wait) {
    for (int i = 0; i < n; ++i) normalise_begin(As[i], st);
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) normalise_finish(As[i], st);
    // The march statistics depend only on the density and the geometry, which a projector keeps:
    // they are counted by its first normalisation and kept.
    bool first = false;
    for (int i = 0; i < n; ++i) first = first || !As[i]->counted;
    if (wait || first) HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i)
        if (!As[i]->counted) read_counters(As[i]), As[i]->counted = true;
}

void normalise_projector(prc_projector* A) { normalise_projectors(&A, 1, build_stream(), true); }

// Host replay of the layout and of every index range the forward and adjoint read (PRC_AUDIT=1):
// the runs of each segment must follow one another from its 64-entry-aligned start, its plane
// sequence and shift list must match the runs (every shift equal to +-D of its ray and step, and
// inside the vectors' margin), and plane reads must stay inside the buffer's one-wave slack and
// shift reads inside the list's slack. Used by the tests before any operator launch; fails
// instead of launching a kernel that could read out of bounds.
void audit_layout(prc_projector* A) {
    const Vol& V = A->hv;
    const size_t units = (size_t)V.ny * V.nz * V.nseg;
    const int R1 = V.n_rays + 1, LP = V.LP;
    std::vector<RowInfo> ri(units * R1), ti(units * R1);
    std::vector<int> kf(R1), pl((size_t)R1 * LP);
    HIPCHK(hipMemcpy(ri.data(), A->rinfo, ri.size() * sizeof(RowInfo), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ti.data(), A->tinfo, ti.size() * sizeof(RowInfo), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(kf.data(), A->kfirst, R1 * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pl.data(), A->plin, pl.size() * sizeof(int), hipMemcpyDeviceToHost));
    const long long shift_max = A->margin - kWave;
    const int P = std::max(PlaneLanes<float>::P, PlaneLanes<__half>::P);
    auto check_copy = [&](const std::vector<RowInfo>& info, const PlaneRun* druns, const int* dlist, long long total,
                          long long list_len, int sign, const char* what) {
        std::vector<PlaneRun> runs(units);
        std::vector<int> list((size_t)list_len);
        HIPCHK(hipMemcpy(runs.data(), druns, units * sizeof(PlaneRun), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(list.data(), dlist, list.size() * sizeof(int), hipMemcpyDeviceToHost));
        long long next = 0, lnext = 0;
        for (size_t u = 0; u < units; ++u) {
            const int row = (int)(u / V.nseg), seg = (int)(u % V.nseg);
            const int w = std::min(kWave, std::max(0, A->h_row_len[row] - kWave * seg));
            const PlaneRun& r = runs[u];
            if (r.off != next || r.dbase != lnext)
                fail(PRC_ERR_INVALID, "audit: %s plane sequence of segment %zu is misplaced", what, u);
            long long off = next;
            int planes = 0;
            for (int p = 0; p < R1; ++p) {
                const RowInfo& e = info[u * R1 + p];
                if (e.off != off || e.len != w || (w > 0 && e.xa != A->h_row_xa[row] + kWave * seg) || e.cnt < 0 ||
                    (w == 0 && e.cnt != 0) || (p > 0 && kf[p] + e.cnt > LP))
                    fail(PRC_ERR_INVALID, "audit: %s layout of segment %zu ray %d is inconsistent", what, u, p);
                for (int k = 0; k < e.cnt; ++k) {
                    const long long d = pl[(size_t)p * LP + kf[p] + k];
                    if (d > shift_max || d < -shift_max)
                        fail(PRC_ERR_INVALID, "audit: %s shift %lld of ray %d exceeds the margin", what, d, p);
                    if (list[(size_t)list_slot(lnext + planes + k, A->wbytes == 2)] != sign * d)
                        fail(PRC_ERR_INVALID, "audit: %s shift list of segment %zu ray %d step %d is wrong", what, u, p, k);
                }
                off += run_size(e.cnt, e.len);
                planes += e.cnt;
            }
            if (r.planes != planes) fail(PRC_ERR_INVALID, "audit: %s segment %zu counts %d planes, not %d", what, u, r.planes, planes);
            for (long long k = planes; k < align_up(planes, kListAlign); ++k)
                if (list[(size_t)list_slot(lnext + k, A->wbytes == 2)] != 0)
                    fail(PRC_ERR_INVALID, "audit: %s list padding is not zero", what);
            // Plane reads: lanes read up to one wave past the last plane's start; shift reads (the
            // last group's prefetch) up to 2P - 2 entries past the list's end.
            if (planes > 0 && (r.off < 0 || r.off + (long long)(planes - 1) * w + kWave - 1 >= total + kWave))
                fail(PRC_ERR_INVALID, "audit: %s planes of segment %zu read outside the buffer", what, u);
            if (lnext + planes + 2 * P - 2 >= list_len + kListSlack)
                fail(PRC_ERR_INVALID, "audit: %s shifts of segment %zu read outside the list", what, u);
            next = align_up(off, kWave);
            lnext += align_up(planes, kListAlign);
        }
        if (next != total || lnext != list_len)
            fail(PRC_ERR_INVALID, "audit: %s layout holds %lld entries and %lld shifts, not %lld and %lld", what, next,
                 lnext, total, list_len);
    };
    const long long* totals = (const long long*)A->h_stage;  // entries and list lengths of this layout
    check_copy(ri, A->sruns, A->slist, A->stats.weight_entries, totals[2], 1, "source-major");
    check_copy(ti, A->truns, A->tlist, A->stats.target_entries, totals[3], -1, "target-major");
    {  // the fill's ray-major copy of the target-major runs
        std::vector<TargetRun> tm(units * R1);
        HIPCHK(hipMemcpy(tm.data(), A->tmaj, tm.size() * sizeof(TargetRun), hipMemcpyDeviceToHost));
        for (size_t u = 0; u < units; ++u)
            for (int p = 0; p < R1; ++p) {
                const TargetRun& m = tm[(size_t)p * units + u];
                const RowInfo& e = ti[u * R1 + p];
                if (m.off != e.off || m.cnt != e.cnt)
                    fail(PRC_ERR_INVALID, "audit: ray-major target run of segment %zu ray %d differs", u, p);
            }
    }
    auto check_item = [&](const std::vector<RowInfo>& info, const int4& it, const char* what) {
        const int row = it.y, x0 = it.z;
        const int seg = (x0 - A->h_row_xa[row]) / kWave;
        if (A->h_row_len[row] <= 0 || (x0 - A->h_row_xa[row]) % kWave || seg < 0 || seg >= V.nseg)
            fail(PRC_ERR_INVALID, "audit: %s item at row %d x %d is not a segment", what, row, x0);
        const RowInfo& e = info[((size_t)row * V.nseg + seg) * R1];
        if (e.xa != x0 || e.len != it.w) fail(PRC_ERR_INVALID, "audit: %s item and segment disagree", what);
    };
    for (const int4& it : A->h_items_fwd)
        if (it.w > 0) check_item(ti, it, "forward");
    for (const int4& it : A->h_items_act)
        if (it.w > 0) check_item(ri, it, "adjoint");
}

void read_counters(prc_projector* A) {
    unsigned long long c[4];
    HIPCHK(hipMemcpy(c, A->counters, sizeof(c), hipMemcpyDeviceToHost));
    A->stats.vrs_near = (long long)c[0];
    A->stats.vrs_far = (long long)c[1];
    A->stats.deposits = (long long)c[2];
    A->stats.nonzero_entries = (long long)c[3];
}

void release(prc_projector* A) {
    if (g_build_stream) (void)hipStreamSynchronize(g_build_stream);
    for (hipStream_t u : A->users) (void)hipStreamSynchronize(u);
    if (A->h_stage) (void)hipHostFree(A->h_stage), A->h_stage = nullptr;
    dfree(A->dv), dfree(A->rho_alloc), dfree(A->table), dfree(A->support), dfree(A->W), dfree(A->row_xa);
    dfree(A->row_len), dfree(A->rinfo), dfree(A->fstep), dfree(A->fsoa), dfree(A->pstep), dfree(A->kfirst), dfree(A->nstep);
    dfree(A->nlen), dfree(A->items_act), dfree(A->items_all), dfree(A->items_fwd), dfree(A->tiles), dfree(A->queue), dfree(A->counters), dfree(A->scratch_in);
    dfree(A->kglob), dfree(A->row_tot), dfree(A->row_base), dfree(A->total), dfree(A->tinfo), dfree(A->plin);
    dfree(A->unit_planes), dfree(A->tile_sums), dfree(A->cnt_t), dfree(A->kt_t), dfree(A->tmaj), dfree(A->list_base), dfree(A->sruns), dfree(A->truns), dfree(A->slist), dfree(A->tlist);
    dfree(A->scratch_adj), dfree(A->cls), dfree(A->near_list), dfree(A->tmpl), dfree(A->tmpl_sum), dfree(A->memo_meta);
    if (A->wraw) (void)hipFree(A->wraw);
    if (A->traw) (void)hipFree(A->traw);
    A->wraw = A->traw = A->wbuf = A->tbuf = nullptr;
}

// Forward/adjoint over one or more volumes sharing a descriptor table and item lists; the
// schedule (enum Sched) is picked per solver or batch by tune_passes.

// `queue` holds kXcds counters in device memory owned by the caller.
// The gathered input is FP32, or FP16 (`half_in`, the mixed-form solver's iterates) with FP16
// weights.
template <class WT, class VT, int GPOL>
void launch_forward_pol(const Vol* dv, const int4* items, int n_items, const VT* in, float* out, int* queue,
                        const PassShape& sh, unsigned nb, hipStream_t s) {
    if (sh.sched == kSchedQueue) k_forward<WT, VT, kSchedQueue, GPOL><<<nb, 256, 0, s>>>(dv, items, n_items, in, out, queue);
    else if (sh.sched == kSchedChunked) k_forward<WT, VT, kSchedChunked, GPOL><<<nb, 256, 0, s>>>(dv, items, n_items, in, out, queue);
    else k_forward<WT, VT, kSchedGrid, GPOL><<<nb, 256, 0, s>>>(dv, items, n_items, in, out, queue);
}
template <class WT, class VT>
void launch_forward(const Vol* dv, const int4* items, int n_items, const void* g, float* out, int* queue,
                    const PassShape& sh, unsigned nb, hipStream_t s) {
    if (sh.gpol) launch_forward_pol<WT, VT, 1>(dv, items, n_items, (const VT*)g, out, queue, sh, nb, s);
    else launch_forward_pol<WT, VT, 0>(dv, items, n_items, (const VT*)g, out, queue, sh, nb, s);
}
template <class WT, class VT, int GPOL>
void launch_adjoint_pol(const Vol* dv, const int4* items, int n_items, const VT* v, float* out, int* queue,
                        const PassShape& sh, unsigned nb, hipStream_t s) {
    if (sh.sched == kSchedQueue) k_adjoint<WT, VT, kSchedQueue, GPOL><<<nb, 256, 0, s>>>(dv, items, n_items, v, out, queue);
    else if (sh.sched == kSchedChunked) k_adjoint<WT, VT, kSchedChunked, GPOL><<<nb, 256, 0, s>>>(dv, items, n_items, v, out, queue);
    else k_adjoint<WT, VT, kSchedGrid, GPOL><<<nb, 256, 0, s>>>(dv, items, n_items, v, out, queue);
}
template <class WT, class VT>
void launch_adjoint(const Vol* dv, const int4* items, int n_items, const void* in, float* out, int* queue,
                    const PassShape& sh, unsigned nb, hipStream_t s) {
    if (sh.gpol) launch_adjoint_pol<WT, VT, 1>(dv, items, n_items, (const VT*)in, out, queue, sh, nb, s);
    else launch_adjoint_pol<WT, VT, 0>(dv, items, n_items, (const VT*)in, out, queue, sh, nb, s);
}

// Persistent grid of the queue schedule: blocks_per_cu blocks (four waves each) per CU. Fewer
// resident waves keep each XCD's concurrent gathers within fewer rows and leave wave slots to the
// kernels of another stream running beside the pass; the tuner picks the count.
unsigned device_cus() {
    static const unsigned cus = [] {
        int dev = 0, n = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        return (unsigned)std::max(n, 1);
    }();
    return cus;
}

unsigned pass_blocks(const PassShape& sh, int n_items) {
    const unsigned all = blocks_for(n_items, kItemsPerBlock);
    return sh.sched == kSchedQueue ? std::min<unsigned>((unsigned)sh.blocks_per_cu * device_cus(), all) : all;
}

void run_forward(const Vol* dv, const int4* items, int n_items, int wbytes, const void* g, bool half_in, float* out,
                 int* queue, const PassShape& sh, hipStream_t s) {
    const unsigned nb = pass_blocks(sh, n_items);
    if (sh.sched == kSchedQueue) k_reset_queue<<<1, 64, 0, s>>>(queue);
    ProfScope ps("forward", s);
    if (wbytes == 4) launch_forward<float, float>(dv, items, n_items, g, out, queue, sh, nb, s);
    else if (half_in) launch_forward<__half, __half>(dv, items, n_items, g, out, queue, sh, nb, s);
    else launch_forward<__half, float>(dv, items, n_items, g, out, queue, sh, nb, s);
    HIPCHK(hipGetLastError());
}

void run_adjoint(const Vol* dv, const int4* items, int n_items, int wbytes, const void* in, bool half_in, float* out,
                 int* queue, const PassShape& sh, hipStream_t s) {
    const unsigned nb = pass_blocks(sh, n_items);
    if (sh.sched == kSchedQueue) k_reset_queue<<<1, 64, 0, s>>>(queue);
    ProfScope ps("adjoint", s);
    if (wbytes == 4) launch_adjoint<float, float>(dv, items, n_items, in, out, queue, sh, nb, s);
    else if (half_in) launch_adjoint<__half, __half>(dv, items, n_items, in, out, queue, sh, nb, s);
    else launch_adjoint<__half, float>(dv, items, n_items, in, out, queue, sh, nb, s);
    HIPCHK(hipGetLastError());
}

void check_geometry(const prc_geometry* g) {
    if (!g) fail(PRC_ERR_INVALID, "null geometry");
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || !(g->sx > 0) || !(g->sy > 0) || !(g->sz > 0) ||
        !std::isfinite(g->sx) || !std::isfinite(g->sy) || !std::isfinite(g->sz) || g->n_rays < 2 || (g->n_rays & 1) ||
        g->n_rays > INT_MAX - 31 || !(g->support_rho >= 0) || !std::isfinite(g->support_rho) ||
        !(g->air_cap_mm > 0) || !std::isfinite(g->air_cap_mm) || g->isotope < 0 || g->isotope >= 8 ||
        (g->precision != PRC_PRECISION_FP32 && g->precision != PRC_PRECISION_MIXED))
        fail(PRC_ERR_INVALID, "Invalid geometry or solver parameters");
    if (g->nx > INT_MAX / g->ny || g->nx * g->ny > INT_MAX / g->nz)
        fail(PRC_ERR_INVALID, "Volume exceeds the INT_MAX voxel index limit");
    if (g->nx + g->ny + g->nz > (1 << 24)) fail(PRC_ERR_INVALID, "grid too long for the ray tables");
}

// The far part of a projector's forward or adjoint on FP32 device vectors (far-marched projectors:
// the stored passes above covered the near block).
void far_march_pass(prc_projector* A, int mode, const float* in, float* out) {
    if (A->wbytes == 2) launch_far_march<__half, float>(A->dv, A, mode, nullptr, in, out, 0);
    else launch_far_march<float, float>(A->dv, A, mode, nullptr, in, out, 0);
}

// Far-marched volumes of a solver or batch: their active segments, tagged with the volume's index
// in the descriptor table.
struct MarchJob {
    const prc_projector* A;
    const int4* items;
    int n;
};

void run_far_march(const Vol* dv, const std::vector<MarchJob>& jobs, int wbytes, int mode, const void* in, bool half_in,
                   float* out, hipStream_t s) {
    for (const MarchJob& j : jobs) {
        if (wbytes == 4) launch_far_march<float, float>(dv, j.A, mode, j.items, in, out, s, j.n);
        else if (half_in) launch_far_march<__half, __half>(dv, j.A, mode, j.items, in, out, s, j.n);
        else launch_far_march<__half, float>(dv, j.A, mode, j.items, in, out, s, j.n);
    }
}

// ---------------------------------------------------------------------------------------
// Solver (prc.cu:659-710), all volumes of a batch at once. Buffers are concatenated voxel
// vectors; sc (device) holds the per-volume scalars.
struct Batch {
    const Vol* dv = nullptr;
    const int4* items_all = nullptr;
    const int4* items_act = nullptr;
    int n_all = 0, n_act = 0, n_vols = 1, wbytes = 4;
    int* queue = nullptr;  // work-queue counters of the projector passes
    PassShape fwd_shape = kShapes[kDefaultShape], adj_shape = kShapes[kDefaultShape];  // chosen by tune_passes
    const int* item_begin = nullptr;  // per-volume ranges of items_all
    const int4* tiles = nullptr;      // RDP update tiles of every volume
    int n_tiles = 0;
    const int* tile_begin = nullptr;  // per-volume ranges of tiles
    float *g = nullptr, *f = nullptr, *prev = nullptr, *y = nullptr, *hprev = nullptr, *est = nullptr,
          *gf = nullptr;
    __half *gh = nullptr, *rh = nullptr;  // mixed form: FP16 copies of the forward and adjoint inputs
    std::vector<MarchJob> march;          // far-marched volumes (near block stored only)
    Scal* sc = nullptr;
    double* partial = nullptr;
    bool accel = false, any_prior = false;
    hipStream_t stream = 0;
};

// prc.cu:666-672: floor on support, then normalise the counts to the supported input sum.
void batch_initialise(const Batch& B) {
    const unsigned nb = blocks_for(B.n_all, kItemsPerBlock);
    ProfScope ps("initialise", B.stream);
    k_initialize<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.f, B.partial);
    k_reduce_volume<<<B.n_vols, 256, 0, B.stream>>>(B.item_begin, B.partial, B.sc, 0);
    k_finish<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.f, nullptr, nullptr, nullptr, 1);
    HIPCHK(hipGetLastError());
}

// One iteration (prc.cu:676-700); the caller rotates the buffers with batch_rotate afterwards.
void batch_iteration(const Batch& B) {
    const unsigned nb = blocks_for(B.n_all, kItemsPerBlock);
    const float* start = B.f;
    // Mixed form: per-item maxima of the gathered vectors go to B.partial, then each volume's
    // scale, then the FP16 copies.
    double* pmax = B.gh ? B.partial : nullptr;
    {
        ProfScope ps("extrapolate", B.stream);
        if (B.accel) {
            k_extrapolate<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.f, B.prev, B.y, B.gf, pmax);
            start = B.y;
        } else {
            k_fscale<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.f, B.gf, pmax);
        }
        if (B.gh) {
            k_reduce_volume<<<B.n_vols, 256, 0, B.stream>>>(B.item_begin, B.partial, B.sc, 4);
            k_half_copy<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.gf, B.gh, 0);
        }
    }
    // The forward covers the active segments only. Voxels outside them are left as they were:
    // k_ratio uses the forward on support voxels only and rewrites every voxel of est (0 off
    // support), so nothing reads them (the operators' forward, which must write every voxel, adds
    // the zero pieces of each row).
    run_forward(B.dv, B.items_act, B.n_act, B.wbytes, B.gh ? (const void*)B.gh : B.gf, B.gh != nullptr, B.est, B.queue,
                B.fwd_shape, B.stream);
    run_far_march(B.dv, B.march, B.wbytes, kMarchFwd, B.gh ? (const void*)B.gh : B.gf, B.gh != nullptr, B.est, B.stream);
    {
        ProfScope ps("ratio", B.stream);
        k_ratio<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.g, B.est, pmax, B.gh != nullptr);
        if (B.rh) {
            k_reduce_volume<<<B.n_vols, 256, 0, B.stream>>>(B.item_begin, B.partial, B.sc, 5);
            k_half_copy<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.est, B.rh, 1);
        }
    }
    run_adjoint(B.dv, B.items_act, B.n_act, B.wbytes, B.rh ? (const void*)B.rh : B.est, B.rh != nullptr, B.prev,
                B.queue, B.adj_shape, B.stream);
    run_far_march(B.dv, B.march, B.wbytes, kMarchAdj, B.rh ? (const void*)B.rh : B.est, B.rh != nullptr, B.prev, B.stream);
    {
        ProfScope ps("update", B.stream);
        k_update<<<B.n_tiles, 256, 0, B.stream>>>(B.dv, B.tiles, B.sc, start, B.prev, B.est, B.partial);
        if (B.any_prior) k_reduce_volume<<<B.n_vols, 256, 0, B.stream>>>(B.tile_begin, B.partial, B.sc, 1);
    }
    if (B.any_prior || B.accel) {
        ProfScope ps("finish", B.stream);
        k_finish<<<nb, 256, 0, B.stream>>>(B.dv, B.items_all, B.n_all, B.sc, B.est, B.accel ? B.y : nullptr,
                                           B.hprev, B.accel ? B.partial : nullptr, B.any_prior ? 1 : 0);
        if (B.accel) k_reduce_volume<<<B.n_vols, 256, 0, B.stream>>>(B.item_begin, B.partial, B.sc, 2);
    }
    HIPCHK(hipGetLastError());
}

// Autotuner of the projector passes: every launch shape of the forward and the adjoint (schedule
// and persistent-grid size) is timed on the solver's or batch's own
// weights and buffers (the forward into the estimate buffer, the adjoint into the work buffer; the
// caller resets both afterwards), and the fastest is kept. Every shape gives the same result bit
// for bit (same per-lane summation order). PRC_TUNE=0 keeps the default shape; PRC_TUNE=<name>
// forces one (names from shape_name; a bare "queue" is the default two blocks per CU).
struct TuneRecord {
    int vols, wbytes, items_fwd, items_adj;
    float ms[2][kShapeCount];  // [forward, adjoint][shape], mean of two launches
    int choice[2];
};
std::vector<TuneRecord> g_tuning;

int shape_by_name(const char* name) {
    for (int k = 0; k < kShapeCount; ++k)
        if (shape_name(kShapes[k]) == name) return k;
    if (!strcmp(name, "queue")) return kDefaultShape;
    return -1;
}

void tune_passes(Batch& B) {
    const char* env = getenv("PRC_TUNE");
    if (env && env[0] == '0') return;
    if (env && *env) {
        const int k = shape_by_name(env);
        if (k < 0) fail(PRC_ERR_INVALID, "PRC_TUNE=%s names no projector-pass shape", env);
        B.fwd_shape = B.adj_shape = kShapes[k];
        return;
    }
    const bool prof = g_prof_on;
    g_prof_on = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    TuneRecord r{B.n_vols, B.wbytes, B.n_act, B.n_act, {}, {kDefaultShape, kDefaultShape}};
    try {
        for (int op = 0; op < 2; ++op)
            for (int k = 0; k < kShapeCount; ++k) {
                auto launch = [&] {
                    if (op == 0)
                        run_forward(B.dv, B.items_act, B.n_act, B.wbytes, B.gh ? (const void*)B.gh : B.gf,
                                    B.gh != nullptr, B.est, B.queue, kShapes[k], B.stream);
                    else
                        run_adjoint(B.dv, B.items_act, B.n_act, B.wbytes, B.rh ? (const void*)B.rh : B.est,
                                    B.rh != nullptr, B.prev, B.queue, kShapes[k], B.stream);
                };
                launch();  // warm-up
                HIPCHK(hipEventRecord(e0, B.stream));
                launch();
                launch();
                HIPCHK(hipEventRecord(e1, B.stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                r.ms[op][k] = ms / 2;
            }
    } catch (...) {
        (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
        g_prof_on = prof;
        throw;
    }
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    g_prof_on = prof;
    for (int op = 0; op < 2; ++op)
        for (int k = 0; k < kShapeCount; ++k)
            if (r.ms[op][k] < r.ms[op][r.choice[op]]) r.choice[op] = k;
    B.fwd_shape = kShapes[r.choice[0]], B.adj_shape = kShapes[r.choice[1]];
    g_tuning.push_back(r);
}

void batch_rotate(Batch& B) {  // prc.cu:697-700
    if (B.accel) {
        std::swap(B.y, B.hprev);
        std::swap(B.prev, B.f);
    }
    std::swap(B.f, B.est);
}

// Host preprocessing of measured data (prc.cu:598-613): clamp, zero off support, FP64 sum.
double preprocess_data(const prc_projector* A, const float* host_g, float* out, float* vmax) {
    double support_sum = 0.0;
    *vmax = 0.0f;
    for (size_t i = 0; i < A->n; ++i) {
        if (!std::isfinite(host_g[i])) fail(PRC_ERR_INVALID, "Non-finite density or activity at voxel %zu", i);
        const float v = std::max(host_g[i], 0.0f);
        out[i] = A->h_support[i] ? v : 0.0f;
        if (A->h_support[i]) support_sum += v, *vmax = std::max(*vmax, v);
    }
    if (support_sum == 0.0)
        fail(PRC_ERR_NO_ACTIVITY, "No activity in voxels with density >= %g: nothing to do", (double)A->g.support_rho);
    return support_sum;
}

Scal initial_scalars(const prc_projector* A, double support_sum, float vmax, const prc_solve_params& p) {
    Scal sc{};
    (void)vmax;
    sc.hscale = sc.rscale = sc.rinv = 1.0f;  // the mixed form sets its scales every iteration
    sc.support_sum = support_sum;
    sc.eps = (float)(1.0e-6 * support_sum / (double)A->stats.n_active);  // prc.cu:668
    sc.lambda = 0.0f, sc.scale = 1.0f, sc.beta = p.beta, sc.gamma = p.gamma;
    sc.accel = p.accelerate ? 1 : 0;
    return sc;
}

void check_params(const prc_solve_params* p) {
    if (!(p->beta >= 0) || !(p->gamma >= 0) || !std::isfinite(p->beta) || !std::isfinite(p->gamma))
        fail(PRC_ERR_INVALID, "Invalid geometry or solver parameters");
}

}  // namespace

// Several volumes corrected together: concatenated vectors, one launch per stage, and the
// iteration loop captured as a HIP graph (cached per iteration count).
struct prc_batch {
    int n = 0;
    std::vector<prc_projector*> A;
    std::vector<Vol> vols;
    std::vector<Scal> sc0;
    size_t total = 0;
    Batch B;
    float* base[5] = {};  // f, prev, y, hprev, est as allocated
    std::vector<void*> owned;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int exec_iterations = -1;
    float* final_f = nullptr;
    Vol* dv = nullptr;
    // Pinned staging of the descriptors and starting scalars that every run uploads, so that the
    // uploads are queued on the batch's stream without waiting for it; `staged` marks when the
    // last upload has read them.
    Vol* h_vols = nullptr;
    Scal* h_sc = nullptr;
    hipEvent_t staged = nullptr;
};

namespace {

void batch_release(prc_batch* b) {
    if (b->B.stream) {
        (void)hipStreamSynchronize(b->B.stream);
        for (prc_projector* A : b->A)
            if (A) A->users.erase(std::remove(A->users.begin(), A->users.end(), b->B.stream), A->users.end());
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->B.stream) (void)hipStreamDestroy(b->B.stream);
    if (b->staged) (void)hipEventDestroy(b->staged);
    if (b->h_vols) (void)hipHostFree(b->h_vols);
    if (b->h_sc) (void)hipHostFree(b->h_sc);
    b->staged = nullptr, b->h_vols = nullptr, b->h_sc = nullptr;
    for (void* p : b->owned) (void)hipFree(p);
    b->owned.clear();
}

// Restore the starting state (estimate = preprocessed data, zero history, initial scalars) and
// refresh the descriptors, which change when a projector is normalised again.
void batch_reset(prc_batch* b) {
    Batch& B = b->B;
    for (int v = 0; v < b->n; ++v) {
        const size_t vb = b->vols[v].vbase;
        b->vols[v] = b->A[v]->hv;
        b->vols[v].vbase = vb;
    }
    HIPCHK(hipEventSynchronize(b->staged));  // the previous upload has read the staging
    std::copy(b->vols.begin(), b->vols.end(), b->h_vols);
    std::copy(b->sc0.begin(), b->sc0.end(), b->h_sc);
    HIPCHK(hipMemcpyAsync(b->dv, b->h_vols, b->n * sizeof(Vol), hipMemcpyHostToDevice, B.stream));
    B.f = b->base[0], B.prev = b->base[1], B.y = b->base[2], B.hprev = b->base[3], B.est = b->base[4];
    HIPCHK(hipMemcpyAsync(B.f, B.g, b->total * sizeof(float), hipMemcpyDeviceToDevice, B.stream));
    for (int i = 1; i < 5; ++i) HIPCHK(hipMemsetAsync(b->base[i], 0, b->total * sizeof(float), B.stream));
    HIPCHK(hipMemcpyAsync(B.sc, b->h_sc, b->n * sizeof(Scal), hipMemcpyHostToDevice, B.stream));
    HIPCHK(hipEventRecord(b->staged, B.stream));
}

void batch_run(prc_batch* b, int iterations) {
    batch_reset(b);
    Batch& B = b->B;
    b->final_f = B.f;
    if (iterations == 0) return;
    batch_initialise(B);
    if (g_prof_on) {  // direct launches so that each kernel can be timed
        Batch C = B;
        for (int it = 0; it < iterations; ++it) {
            batch_iteration(C);
            batch_rotate(C);
        }
        b->final_f = C.f;
        return;
    }
    if (b->exec_iterations != iterations) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr, b->graph = nullptr, b->exec_iterations = -1;
        HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        Batch C = B;
        try {
            for (int it = 0; it < iterations; ++it) {
                batch_iteration(C);
                batch_rotate(C);
            }
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(B.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(B.stream, &b->graph));
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->exec_iterations = iterations;
        b->final_f = C.f;
    } else {
        Batch C = B;
        for (int it = 0; it < iterations; ++it) batch_rotate(C);
        b->final_f = C.f;
    }
    HIPCHK(hipGraphLaunch(b->exec, B.stream));
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Forward without allocation (prc_projector_forward_noalloc): the margin-padded staging vector
// g = f / W lives in this file-scope array instead of a buffer the library allocates on first
// use. Each GPU holds its own copy until the process ends. Host code never uses the array's name
// as an address: scratch_base() looks the device address up once per device.
constexpr size_t kScratchFloats = size_t(32) << 20;  // 128 MiB: n + 2 margins of a 344x344x127 volume
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

float* scratch_base() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(PRC_ERR_DEVICE, "device %d beyond the library's %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (float*)p;
    }
    return g_scratch[dev];
}

__global__ void k_fill(float* __restrict__ p, long long n, float v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

// =======================================================================================
extern "C" {

const char* prc_version(void) { return "prc-mi355x 0.2 (gfx950)"; }
const char* prc_last_error(void) { return g_error.c_str(); }
int prc_isotope_count(void) { return 8; }
const char* prc_isotope_name(int index) { return index >= 0 && index < 8 ? kIsotopes[index].name : nullptr; }
int prc_isotope_find(const char* name) {
    if (!name) return -1;
    for (int i = 0; i < 8; ++i)
        if (!strcmp(name, kIsotopes[i].name)) return i;
    return -1;
}
float prc_range_cutoff(int isotope, float fraction) {
    if (isotope < 0 || isotope >= 8) return NAN;
    return range_cutoff_host(renormalised(isotope), fraction);
}

int prc_ray_set(int n, float* host_start, float* host_dir) {
    return guarded([&] {
        if (n < 2 || (n & 1) || n > (1 << 26) || !host_start || !host_dir) fail(PRC_ERR_INVALID, "invalid argument");
        float* d = dalloc<float>((size_t)6 * n);
        std::vector<float> h((size_t)6 * n);
        try {
            k_ray_set<<<blocks_for(n, 256), 256>>>(n, d);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
        for (size_t k = 0; k < (size_t)n; ++k)
            for (int c = 0; c < 3; ++c) host_start[3 * k + c] = h[6 * k + c], host_dir[3 * k + c] = h[6 * k + 3 + c];
    });
}

int prc_projector_create(const prc_geometry* geometry, const float* host_rho, const float* host_table,
                         prc_projector** out) {
    return guarded([&] {
        if (!out || !host_rho) fail(PRC_ERR_INVALID, "null argument");
        *out = nullptr;
        check_geometry(geometry);
        const size_t n = (size_t)geometry->nx * geometry->ny * geometry->nz;
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(host_rho[i])) fail(PRC_ERR_INVALID, "Non-finite density at voxel %zu", i);
        prc_projector* A = new prc_projector;
        A->g = *geometry;
        try {
            const double t0 = now_s();
            setup_projector(A, host_rho, host_table);
            HIPCHK(hipDeviceSynchronize());  // setup ran on the null stream
            normalise_projector(A);
            HIPCHK(hipDeviceSynchronize());
            A->stats.norm_seconds = now_s() - t0;
            const char* audit = getenv("PRC_AUDIT");
            if (audit && *audit == '1') audit_layout(A);
        } catch (...) {
            release(A);
            delete A;
            throw;
        }
        *out = A;
    });
}

int prc_projector_normalise(prc_projector* A) {
    return guarded([&] {
        if (!A) fail(PRC_ERR_INVALID, "null argument");
        normalise_projector(A);
    });
}

int prc_projectors_normalise(int n, prc_projector* const* projectors) {
    return guarded([&] {
        if (n <= 0 || !projectors) fail(PRC_ERR_INVALID, "invalid argument");
        for (int i = 0; i < n; ++i) {
            if (!projectors[i]) fail(PRC_ERR_INVALID, "null projector");
            for (int j = 0; j < i; ++j)
                if (projectors[j] == projectors[i]) fail(PRC_ERR_INVALID, "a projector is listed twice");
        }
        normalise_projectors(projectors, n, build_stream(), true);
    });
}

void prc_projector_destroy(prc_projector* A) {
    if (!A) return;
    release(A);
    delete A;
}

int prc_projector_forward(prc_projector* A, const float* dev_in, float* dev_out) {
    return guarded([&] {
        if (!A || !dev_in || !dev_out) fail(PRC_ERR_INVALID, "null argument");
        if (!A->scratch_in) {
            A->scratch_in = dalloc<float>(A->n + 2 * A->margin);
            HIPCHK(hipMemset(A->scratch_in, 0, (A->n + 2 * A->margin) * sizeof(float)));
        }
        float* g = A->scratch_in + A->margin;  // f / W with the forward's gather margin
        {
            ProfScope ps("extrapolate", 0);
            k_fscale<<<blocks_for(A->n_all, kItemsPerBlock), 256>>>(A->dv, A->items_all, A->n_all, dev_in, g, nullptr);
            HIPCHK(hipGetLastError());
        }
        run_forward(A->dv, A->items_fwd, A->n_fwd, A->wbytes, g, false, dev_out, A->queue, A->fwd_shape, 0);
        if (A->far_marched) far_march_pass(A, kMarchFwd, g, dev_out);
    });
}

int prc_projector_forward_noalloc(prc_projector* A, const float* dev_in, float* dev_out) {
    return guarded([&] {
        if (!A || !dev_in || !dev_out) fail(PRC_ERR_INVALID, "null argument");
        const size_t margin = (size_t)A->margin, need = A->n + 2 * margin;
        if (need > kScratchFloats)
            fail(PRC_ERR_INVALID,
                 "the no-allocation forward holds %zu floats of scratch; this volume needs %zu (voxels plus two "
                 "gather margins): use prc_projector_forward",
                 kScratchFloats, need);
        float* base = scratch_base();
        // The margins are read by the gathers (against zero weights): clear them, with a kernel so
        // that the call can be captured in a graph. The volume part is written whole by k_fscale.
        const unsigned fb = std::min<unsigned>(blocks_for((long long)margin, 256), 1024u);
        k_fill<<<fb, 256>>>(base, (long long)margin, 0.0f);
        k_fill<<<fb, 256>>>(base + margin + A->n, (long long)margin, 0.0f);
        float* g = base + margin;
        k_fscale<<<blocks_for(A->n_all, kItemsPerBlock), 256>>>(A->dv, A->items_all, A->n_all, dev_in, g, nullptr);
        HIPCHK(hipGetLastError());
        run_forward(A->dv, A->items_fwd, A->n_fwd, A->wbytes, g, false, dev_out, A->queue, A->fwd_shape, 0);
        if (A->far_marched) far_march_pass(A, kMarchFwd, g, dev_out);
    });
}

int prc_projector_adjoint(prc_projector* A, const float* dev_in, float* dev_out) {
    return guarded([&] {
        if (!A || !dev_in || !dev_out) fail(PRC_ERR_INVALID, "null argument");
        if (!A->scratch_adj) {
            A->scratch_adj = dalloc<float>(A->n + 2 * A->margin);
            HIPCHK(hipMemset(A->scratch_adj, 0, (A->n + 2 * A->margin) * sizeof(float)));
        }
        float* in = A->scratch_adj + A->margin;  // the input with the adjoint's gather margin
        HIPCHK(hipMemcpyAsync(in, dev_in, A->n * sizeof(float), hipMemcpyDeviceToDevice, 0));
        run_adjoint(A->dv, A->items_act, A->n_act, A->wbytes, in, false, dev_out, A->queue, A->adj_shape, 0);
        if (A->far_marched) far_march_pass(A, kMarchAdj, in, dev_out);
    });
}

static int host_op(prc_projector* A, const float* in, float* out, bool fwd) {
    return guarded([&] {
        if (!A || !in || !out) fail(PRC_ERR_INVALID, "null argument");
        float* din = dalloc<float>(A->n);
        float* dout = nullptr;
        try {
            dout = dalloc<float>(A->n);
            HIPCHK(hipMemcpy(din, in, A->n * sizeof(float), hipMemcpyHostToDevice));
            if (fwd) {
                const int rc = prc_projector_forward(A, din, dout);
                if (rc) throw Failure{rc};
            } else {
                HIPCHK(hipMemcpy(dout, out, A->n * sizeof(float), hipMemcpyHostToDevice));
                const int rc = prc_projector_adjoint(A, din, dout);
                if (rc) throw Failure{rc};
            }
            HIPCHK(hipMemcpy(out, dout, A->n * sizeof(float), hipMemcpyDeviceToHost));
        } catch (...) {
            dfree(din), dfree(dout);
            throw;
        }
        dfree(din), dfree(dout);
    });
}

int prc_projector_forward_host(prc_projector* A, const float* in, float* out) { return host_op(A, in, out, true); }
int prc_projector_adjoint_host(prc_projector* A, const float* in, float* out) { return host_op(A, in, out, false); }

int prc_projector_norm(prc_projector* A, float* host_w) {
    return guarded([&] {
        if (!A || !host_w) fail(PRC_ERR_INVALID, "null argument");
        HIPCHK(hipMemcpy(host_w, A->W, A->n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

const float* prc_projector_norm_device(prc_projector* A) { return A ? A->W : nullptr; }

int prc_projector_table(prc_projector* A, float* host_table, float* scale) {
    return guarded([&] {
        if (!A || !host_table) fail(PRC_ERR_INVALID, "null argument");
        HIPCHK(hipMemcpy(host_table, A->table, (kBins + 1) * sizeof(float), hipMemcpyDeviceToHost));
        if (scale) *scale = A->hv.scale;
    });
}

int prc_projector_stats(prc_projector* A, prc_stats* stats) {
    return guarded([&] {
        if (!A || !stats) fail(PRC_ERR_INVALID, "null argument");
        *stats = A->stats;
    });
}

static Batch solver_batch(prc_solver* S);

int prc_solver_create(prc_projector* A, const float* host_g, const prc_solve_params* params, prc_solver** out) {
    return guarded([&] {
        if (!A || !host_g || !params || !out) fail(PRC_ERR_INVALID, "null argument");
        *out = nullptr;
        check_params(params);
        const size_t n = A->n;
        std::vector<float> g(n);
        float vmax = 0.0f;
        const double support_sum = preprocess_data(A, host_g, g.data(), &vmax);
        prc_solver* S = new prc_solver;
        S->A = A;
        S->p = *params;
        S->support_sum = support_sum;
        S->sc0 = initial_scalars(A, support_sum, vmax, *params);
        try {
            const size_t padded = n + 2 * (size_t)A->margin;
            for (int i = 0; i < 7; ++i) {
                S->raw[i] = dalloc<float>(padded);
                HIPCHK(hipMemset(S->raw[i], 0, padded * sizeof(float)));
            }
            S->g = S->raw[5] + A->margin, S->gf = S->raw[6] + A->margin;
            if (A->wbytes == 2)
                for (int i = 0; i < 2; ++i) {
                    S->hraw[i] = dalloc<__half>(padded);
                    HIPCHK(hipMemset(S->hraw[i], 0, padded * sizeof(__half)));
                }
            for (int i = 0; i < 5; ++i) S->base[i] = S->raw[i] + A->margin;
            HIPCHK(hipMemcpy(S->g, g.data(), n * sizeof(float), hipMemcpyHostToDevice));
            S->sc = dalloc<Scal>(1);
            S->partial = dalloc<double>(std::max<size_t>(2 * (size_t)A->n_all, (size_t)A->n_tiles));
            int ib[2] = {0, A->n_all}, tb[2] = {0, A->n_tiles};
            S->item_begin = dalloc<int>(2);
            HIPCHK(hipMemcpy(S->item_begin, ib, sizeof(ib), hipMemcpyHostToDevice));
            S->tile_begin = dalloc<int>(2);
            HIPCHK(hipMemcpy(S->tile_begin, tb, sizeof(tb), hipMemcpyHostToDevice));
            int rc = prc_solver_reset(S);
            if (rc) throw Failure{rc};
            if (!A->tuned) {  // the projector's first solver tunes its passes
                Batch B = solver_batch(S);
                tune_passes(B);
                A->fwd_shape = B.fwd_shape, A->adj_shape = B.adj_shape, A->tuned = true;
                rc = prc_solver_reset(S);
                if (rc) throw Failure{rc};
            }
        } catch (...) {
            prc_solver_destroy(S);
            throw;
        }
        *out = S;
    });
}

int prc_solver_reset(prc_solver* S) {
    return guarded([&] {
        if (!S) fail(PRC_ERR_INVALID, "null argument");
        const size_t n = S->A->n;
        S->f = S->base[0], S->prev = S->base[1], S->y = S->base[2], S->hprev = S->base[3], S->est = S->base[4];
        HIPCHK(hipMemcpyAsync(S->f, S->g, n * sizeof(float), hipMemcpyDeviceToDevice, 0));
        for (int i = 1; i < 5; ++i) HIPCHK(hipMemsetAsync(S->base[i], 0, n * sizeof(float), 0));
        HIPCHK(hipMemcpy(S->sc, &S->sc0, sizeof(Scal), hipMemcpyHostToDevice));
        S->initialised = false;
    });
}

void prc_solver_destroy(prc_solver* S) {
    if (!S) return;
    for (float*& b : S->raw) dfree(b);
    for (__half*& b : S->hraw) dfree(b);
    dfree(S->sc), dfree(S->partial), dfree(S->item_begin), dfree(S->tile_begin);
    delete S;
}

static Batch solver_batch(prc_solver* S) {
    Batch B;
    prc_projector* A = S->A;
    B.dv = A->dv, B.items_all = A->items_all, B.items_act = A->items_act, B.n_all = A->n_all, B.n_act = A->n_act;
    B.queue = A->queue;
    B.fwd_shape = A->fwd_shape, B.adj_shape = A->adj_shape;
    B.n_vols = 1, B.wbytes = A->wbytes, B.item_begin = S->item_begin;
    B.tiles = A->tiles, B.n_tiles = A->n_tiles, B.tile_begin = S->tile_begin;
    if (A->far_marched) B.march.push_back(MarchJob{A, A->items_act, A->n_act});
    B.g = S->g, B.f = S->f, B.prev = S->prev, B.y = S->y, B.hprev = S->hprev, B.est = S->est, B.gf = S->gf;
    if (S->hraw[0]) B.gh = S->hraw[0] + S->A->margin, B.rh = S->hraw[1] + S->A->margin;
    B.sc = S->sc, B.partial = S->partial;
    B.accel = S->p.accelerate != 0, B.any_prior = S->p.beta > 0.0f;
    return B;
}

int prc_solver_iterate(prc_solver* S, int n) {
    return guarded([&] {
        if (!S || n < 0) fail(PRC_ERR_INVALID, "invalid argument");
        if (n == 0) return;
        Batch B = solver_batch(S);
        if (!S->initialised) {
            batch_initialise(B);
            S->initialised = true;
        }
        for (int it = 0; it < n; ++it) {
            batch_iteration(B);
            batch_rotate(B);
        }
        S->f = B.f, S->prev = B.prev, S->y = B.y, S->hprev = B.hprev, S->est = B.est;
    });
}

int prc_solver_sync(prc_solver* S) {
    return guarded([&] {
        if (!S) fail(PRC_ERR_INVALID, "null argument");
        HIPCHK(hipDeviceSynchronize());
    });
}

int prc_solver_estimate(prc_solver* S, float* host_out) {
    return guarded([&] {
        if (!S || !host_out) fail(PRC_ERR_INVALID, "null argument");
        HIPCHK(hipMemcpy(host_out, S->f, S->A->n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

const float* prc_solver_estimate_device(prc_solver* S) { return S ? S->f : nullptr; }

int prc_solver_reproject(prc_solver* S, float* host_out) {
    return guarded([&] {
        if (!S || !host_out) fail(PRC_ERR_INVALID, "null argument");
        const int rc = prc_projector_forward(S->A, S->f, S->est);
        if (rc) throw Failure{rc};
        HIPCHK(hipMemcpy(host_out, S->est, S->A->n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int prc_solver_sums(prc_solver* S, double* estimate_sum, double* support_sum) {
    return guarded([&] {
        if (!S) fail(PRC_ERR_INVALID, "null argument");
        prc_projector* A = S->A;
        k_sum_partial<<<blocks_for(A->n_all, kItemsPerBlock), 256>>>(A->dv, A->items_all, A->n_all, S->f, S->partial);
        k_reduce_volume<<<1, 256>>>(S->item_begin, S->partial, S->sc, 3);
        HIPCHK(hipGetLastError());
        Scal sc;
        HIPCHK(hipMemcpy(&sc, S->sc, sizeof(Scal), hipMemcpyDeviceToHost));
        if (estimate_sum) *estimate_sum = sc.sum;
        if (support_sum) *support_sum = S->support_sum;
    });
}

int prc_batch_create(int n, prc_projector* const* projectors, const float* const* host_g,
                     const prc_solve_params* params, prc_batch** out) {
    return guarded([&] {
        if (n <= 0 || !projectors || !host_g || !params || !out) fail(PRC_ERR_INVALID, "invalid argument");
        *out = nullptr;
        prc_batch* b = new prc_batch;
        try {
            b->n = n;
            b->A.assign(projectors, projectors + n);
            b->vols.resize(n);
            b->sc0.resize(n);
            std::vector<int4> all, act, tiles;
            std::vector<int> begin(n + 1, 0), tbegin(n + 1, 0);
            const int wbytes = projectors[0] ? projectors[0]->wbytes : 4;
            const bool accel = params[0].accelerate != 0;
            bool prior = false;
            for (int v = 0; v < n; ++v) {
                prc_projector* A = projectors[v];
                if (!A) fail(PRC_ERR_INVALID, "null projector");
                if (A->wbytes != wbytes) fail(PRC_ERR_INVALID, "a batch must share one precision");
                if ((params[v].accelerate != 0) != accel) fail(PRC_ERR_INVALID, "a batch must share the acceleration flag");
                check_params(&params[v]);
                prior = prior || params[v].beta > 0.0f;
                b->vols[v] = A->hv;
                b->vols[v].vbase = b->total + A->margin;  // each volume keeps its gather margins
                begin[v] = (int)all.size();
                tbegin[v] = (int)tiles.size();
                for (const int4& t : A->h_tiles) tiles.push_back(make_int4(v, t.y, t.z, t.w));
                for (const int4& it : A->h_items_all) all.push_back(make_int4(v, it.y, it.z, it.w));
                for (const int4& it : A->h_items_act) act.push_back(make_int4(v, it.y, it.z, it.w));
                b->total += A->n + 2 * (size_t)A->margin;
            }
            begin[n] = (int)all.size();
            tbegin[n] = (int)tiles.size();
            std::vector<float> g(b->total, 0.0f);
            for (int v = 0; v < n; ++v) {
                float vmax = 0.0f;
                const double ss = preprocess_data(projectors[v], host_g[v], g.data() + b->vols[v].vbase, &vmax);
                b->sc0[v] = initial_scalars(projectors[v], ss, vmax, params[v]);
            }
            auto own = [&](void* p) {
                b->owned.push_back(p);
                return p;
            };
            Batch& B = b->B;
            b->dv = (Vol*)own(dalloc<Vol>(n));
            int4* d_all = (int4*)own(dalloc<int4>(all.size()));
            int4* d_act = (int4*)own(dalloc<int4>(std::max<size_t>(act.size(), 1)));
            int* d_begin = (int*)own(dalloc<int>(n + 1));
            HIPCHK(hipMemcpy(d_all, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_act, act.data(), act.size() * sizeof(int4), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_begin, begin.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
            int4* d_tiles = (int4*)own(dalloc<int4>(tiles.size()));
            int* d_tbegin = (int*)own(dalloc<int>(n + 1));
            HIPCHK(hipMemcpy(d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_tbegin, tbegin.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
            B.tiles = d_tiles, B.n_tiles = (int)tiles.size(), B.tile_begin = d_tbegin;
            for (int v = 0; v < n; ++v) {  // far-marched volumes: their own active items, tagged v
                prc_projector* A = projectors[v];
                if (!A->far_marched) continue;
                std::vector<int4> mine;
                for (const int4& it : A->h_items_act) mine.push_back(make_int4(v, it.y, it.z, it.w));
                int4* d = (int4*)own(dalloc<int4>(mine.size()));
                HIPCHK(hipMemcpy(d, mine.data(), mine.size() * sizeof(int4), hipMemcpyHostToDevice));
                B.march.push_back(MarchJob{A, d, (int)mine.size()});
            }
            B.dv = b->dv, B.items_all = d_all, B.items_act = d_act, B.n_all = (int)all.size(), B.n_act = (int)act.size();
            B.queue = (int*)own(dalloc<int>(kXcds));
            B.n_vols = n, B.wbytes = wbytes, B.item_begin = d_begin, B.accel = accel, B.any_prior = prior;
            B.g = (float*)own(dalloc<float>(b->total));
            B.gf = (float*)own(dalloc<float>(b->total));
            HIPCHK(hipMemset(B.gf, 0, b->total * sizeof(float)));
            if (wbytes == 2) {
                B.gh = (__half*)own(dalloc<__half>(b->total));
                B.rh = (__half*)own(dalloc<__half>(b->total));
                HIPCHK(hipMemset(B.gh, 0, b->total * sizeof(__half)));
                HIPCHK(hipMemset(B.rh, 0, b->total * sizeof(__half)));
            }
            for (int i = 0; i < 5; ++i) b->base[i] = (float*)own(dalloc<float>(b->total));
            HIPCHK(hipMemcpy(B.g, g.data(), b->total * sizeof(float), hipMemcpyHostToDevice));
            B.sc = (Scal*)own(dalloc<Scal>(n));
            B.partial = (double*)own(dalloc<double>(std::max(2 * all.size(), tiles.size())));
            HIPCHK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
            HIPCHK(hipEventRecord(b->staged, B.stream));
            HIPCHK(hipHostMalloc((void**)&b->h_vols, n * sizeof(Vol)));
            HIPCHK(hipHostMalloc((void**)&b->h_sc, n * sizeof(Scal)));
            for (int v = 0; v < n; ++v) projectors[v]->users.push_back(B.stream);
            batch_reset(b);
            tune_passes(B);
            HIPCHK(hipStreamSynchronize(B.stream));
        } catch (...) {
            batch_release(b);
            delete b;
            throw;
        }
        *out = b;
    });
}

int prc_batch_normalise(prc_batch* b) {
    return guarded([&] {
        if (!b) fail(PRC_ERR_INVALID, "null argument");
        normalise_projectors(b->A.data(), b->n, b->B.stream, false);
    });
}

int prc_batch_run(prc_batch* b, int iterations) {
    return guarded([&] {
        if (!b || iterations < 0) fail(PRC_ERR_INVALID, "invalid argument");
        batch_run(b, iterations);
    });
}

int prc_batch_sync(prc_batch* b) {
    return guarded([&] {
        if (!b) fail(PRC_ERR_INVALID, "null argument");
        HIPCHK(hipStreamSynchronize(b->B.stream));
    });
}

int prc_batch_estimate(prc_batch* b, int v, float* host_out) {
    return guarded([&] {
        if (!b || !host_out || v < 0 || v >= b->n || !b->final_f) fail(PRC_ERR_INVALID, "invalid argument");
        HIPCHK(hipStreamSynchronize(b->B.stream));
        HIPCHK(hipMemcpy(host_out, b->final_f + b->vols[v].vbase, b->A[v]->n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

void prc_batch_destroy(prc_batch* b) {
    if (!b) return;
    batch_release(b);
    delete b;
}

int prc_batch_correct(int n, prc_projector* const* projectors, const float* const* host_g,
                      const prc_solve_params* params, int iterations, float* const* host_out) {
    prc_batch* b = nullptr;
    int rc = prc_batch_create(n, projectors, host_g, params, &b);
    if (rc) return rc;
    rc = prc_batch_run(b, iterations);
    for (int v = 0; v < n && rc == PRC_OK; ++v) rc = prc_batch_estimate(b, v, host_out[v]);
    prc_batch_destroy(b);
    return rc;
}

int prc_tuning_report(char* json, size_t capacity) {
    return guarded([&] {
        std::string s = "[";
        for (size_t i = 0; i < g_tuning.size(); ++i) {
            const TuneRecord& r = g_tuning[i];
            char buf[256];
            snprintf(buf, sizeof(buf), "%s{\"volumes\": %d, \"weight_bytes\": %d, \"forward_items\": %d, \"adjoint_items\": %d",
                     i ? ", " : "", r.vols, r.wbytes, r.items_fwd, r.items_adj);
            s += buf;
            for (int op = 0; op < 2; ++op) {
                s += op ? ", \"adjoint_ms\": {" : ", \"forward_ms\": {";
                for (int k = 0; k < kShapeCount; ++k) {
                    snprintf(buf, sizeof(buf), "%s\"%s\": %.4f", k ? ", " : "", shape_name(kShapes[k]).c_str(), r.ms[op][k]);
                    s += buf;
                }
                s += "}";
            }
            s += ", \"forward\": \"" + shape_name(kShapes[r.choice[0]]) + "\", \"adjoint\": \"" +
                 shape_name(kShapes[r.choice[1]]) + "\"}";
        }
        s += "]";
        if (!json || capacity < s.size() + 1) fail(PRC_ERR_INVALID, "report needs %zu bytes", s.size() + 1);
        memcpy(json, s.c_str(), s.size() + 1);
    });
}

int prc_profile_enable(int on) {
    return guarded([&] {
        HIPCHK(hipDeviceSynchronize());
        g_prof_on = on != 0;
    });
}

int prc_profile_report(char* json, size_t capacity) {
    return guarded([&] {
        HIPCHK(hipDeviceSynchronize());
        struct Agg {
            long long count = 0;
            double total = 0.0, lo = 1e300, hi = 0.0;
        };
        std::vector<std::pair<std::string, Agg>> agg;
        for (const ProfRecord& r : g_prof) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, r.begin, r.end));
            auto it = std::find_if(agg.begin(), agg.end(), [&](const auto& e) { return e.first == r.name; });
            if (it == agg.end()) agg.emplace_back(r.name, Agg{}), it = agg.end() - 1;
            it->second.count++, it->second.total += ms;
            it->second.lo = std::min(it->second.lo, (double)ms), it->second.hi = std::max(it->second.hi, (double)ms);
        }
        for (const ProfRecord& r : g_prof) (void)hipEventDestroy(r.begin), (void)hipEventDestroy(r.end);
        g_prof.clear();
        std::string s = "{";
        for (size_t i = 0; i < agg.size(); ++i) {
            char buf[256];
            snprintf(buf, sizeof(buf), "%s\"%s\": {\"count\": %lld, \"total_ms\": %.6f, \"mean_ms\": %.6f, \"min_ms\": %.6f, \"max_ms\": %.6f}",
                     i ? ", " : "", agg[i].first.c_str(), agg[i].second.count, agg[i].second.total,
                     agg[i].second.total / agg[i].second.count, agg[i].second.lo, agg[i].second.hi);
            s += buf;
        }
        s += "}";
        if (!json || capacity < s.size() + 1) fail(PRC_ERR_INVALID, "report needs %zu bytes", s.size() + 1);
        memcpy(json, s.c_str(), s.size() + 1);
    });
}

}  // extern "C"
