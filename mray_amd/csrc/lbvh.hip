// LBVH build on gfx950: primitive centres/boxes, scene box, 63-bit Morton
// codes, LDS-staged stable LSD radix sort, Karras hierarchy and bottom-up
// refit.  Reference: Tracer/AcceleratorLBVH.{hpp,cu}, Tracer/AcceleratorWork.kt.h.
#include "common.h"
#include "lbvh.h"

#include <mutex>
#include <string>

namespace mray
{

// ---------------------------------------------------------------------------
// KCGenPrimCenters + KCGeneratePrimAABBs (AcceleratorWork.kt.h:5-123):
// one thread per triangle.  Centre = p0*k + p1*k + p2*k, k = 0.333333333
// (PrimitiveDefaultTriangle.hpp:107-115), box = component min/max.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kTPB)
KCPrimCentersAabbs(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                   uint32_t n, float* __restrict__ centers, float* __restrict__ aabbs)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n) return;
    uint32_t i0 = idx[3 * t + 0], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
    V3 p0 = {pos[3 * i0], pos[3 * i0 + 1], pos[3 * i0 + 2]};
    V3 p1 = {pos[3 * i1], pos[3 * i1 + 1], pos[3 * i1 + 2]};
    V3 p2 = {pos[3 * i2], pos[3 * i2 + 1], pos[3 * i2 + 2]};
    const float k = 0.333333333f;
    V3 c = (p0 * k + p1 * k) + p2 * k;
    if(centers)
    {
        centers[3 * t + 0] = c.x;
        centers[3 * t + 1] = c.y;
        centers[3 * t + 2] = c.z;
    }
    if(aabbs)
    {
        float* b = aabbs + 6 * size_t(t);
        b[0] = MinF(MinF(p0.x, p1.x), p2.x);
        b[1] = MinF(MinF(p0.y, p1.y), p2.y);
        b[2] = MinF(MinF(p0.z, p1.z), p2.z);
        b[3] = MaxF(MaxF(p0.x, p1.x), p2.x);
        b[4] = MaxF(MaxF(p0.y, p1.y), p2.y);
        b[5] = MaxF(MaxF(p0.z, p1.z), p2.z);
    }
}

// ---------------------------------------------------------------------------
// Scene box: two-level min/max reduction (exact, order-free).
// ---------------------------------------------------------------------------
static constexpr uint32_t kReduceBlocks = 1024;

__global__ void __launch_bounds__(kTPB)
KCReduceAabbPartial(const float* __restrict__ aabbs, uint32_t n, float* __restrict__ partial)
{
    __shared__ float s[6][kTPB];
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const float* a = aabbs + 6 * size_t(i);
        for(int k = 0; k < 3; k++)
        {
            b[k] = MinF(b[k], a[k]);
            b[3 + k] = MaxF(b[3 + k], a[3 + k]);
        }
    }
    for(int k = 0; k < 6; k++) s[k][threadIdx.x] = b[k];
    __syncthreads();
    for(uint32_t st = blockDim.x / 2; st > 0; st >>= 1)
    {
        if(threadIdx.x < st)
        {
            for(int k = 0; k < 3; k++)
            {
                s[k][threadIdx.x] = MinF(s[k][threadIdx.x], s[k][threadIdx.x + st]);
                s[3 + k][threadIdx.x] = MaxF(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
            }
        }
        __syncthreads();
    }
    if(threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// KCGenMortonCode (AcceleratorLBVH.cu:96-168), bit-exact: float offset from
// the box minimum clamped at 0, double product with 2^21 / maxSide, rounded
// back to float, lroundf, clamp to [0, 2^21 - 1], Compose3D<uint64_t>.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t Expand3D64(uint32_t v)
{
    uint64_t x = v;
    x &= 0x1fffff;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
__device__ __forceinline__ uint32_t Expand3D32(uint32_t x)
{
    x &= 0x000003ff;
    x = (x ^ (x << 16)) & 0xff0000ff;
    x = (x ^ (x << 8)) & 0x0300f00f;
    x = (x ^ (x << 4)) & 0x030c30c3;
    x = (x ^ (x << 2)) & 0x09249249;
    return x;
}

template<uint32_t BITS>
__device__ __forceinline__ void Quantize(const float* __restrict__ c, const float* __restrict__ box,
                                         uint32_t q[3])
{
    float sx = box[3] - box[0], sy = box[4] - box[1], sz = box[5] - box[2];
    float maxSide = sx;
    if(sy > maxSide) maxSide = sy;
    if(sz > maxSide) maxSide = sz;
    const double sliceCount = double(1ull << BITS);
    double deltaRecip = sliceCount / double(maxSide);
    const uint32_t last = (1u << BITS) - 1u;
    for(int k = 0; k < 3; k++)
    {
        float d = MaxF(c[k] - box[k], 0.0f);
        float qf = float(double(d) * deltaRecip);
        uint32_t qi = uint32_t(int32_t(lroundf(qf)));
        q[k] = (qi > last) ? last : qi;
    }
}

__global__ void __launch_bounds__(kTPB)
KCMorton63(const float* __restrict__ centers, uint32_t n, const float* __restrict__ box,
           uint64_t* __restrict__ codes)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float c[3] = {centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]};
    float b[6];
    for(int k = 0; k < 6; k++) b[k] = box[k];
    uint32_t q[3];
    Quantize<21>(c, b, q);
    codes[i] = Expand3D64(q[0]) | (Expand3D64(q[1]) << 1) | (Expand3D64(q[2]) << 2);
}

__global__ void __launch_bounds__(kTPB)
KCMorton30(const float* __restrict__ centers, uint32_t n, const float* __restrict__ box,
           uint32_t* __restrict__ codes)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float c[3] = {centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]};
    float b[6];
    for(int k = 0; k < 6; k++) b[k] = box[k];
    uint32_t q[3];
    Quantize<10>(c, b, q);
    codes[i] = Expand3D32(q[0]) | (Expand3D32(q[1]) << 1) | (Expand3D32(q[2]) << 2);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (u64 key, u32 value), 8 bits per pass.
// Per pass: tile histograms -> one exclusive scan (digit-major) -> stable
// scatter.  The scatter ranks a 4096-key tile in 16 rounds of 256 keys; inside
// a round each wave64 splits its keys by digit with 8 ballots (peers of a lane
// = lanes holding the same digit), and the four waves are merged through a
// [4][256] LDS count table, so equal digits keep input order.
// ---------------------------------------------------------------------------
static constexpr uint32_t kRsThreads = 256;
static constexpr uint32_t kRsRounds = 16;
static constexpr uint32_t kRsTile = kRsThreads * kRsRounds;
static constexpr uint32_t kRsWaves = kRsThreads / 64;

__global__ void __launch_bounds__(kRsThreads)
KCRadixHist(const uint64_t* __restrict__ keys, uint32_t n, uint32_t shift, uint32_t numTiles,
            uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kRsTile;
    for(uint32_t r = 0; r < kRsRounds; r++)
    {
        uint32_t i = base + r * kRsThreads + threadIdx.x;
        if(i < n) atomicAdd(&h[uint32_t(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[size_t(threadIdx.x) * numTiles + blockIdx.x] = h[threadIdx.x];
}

// Single-workgroup exclusive scan of `count` u32 values in place.
__global__ void __launch_bounds__(1024)
KCExclusiveScanSingle(uint32_t* __restrict__ data, uint32_t count, uint32_t* __restrict__ total)
{
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    if(threadIdx.x == 0) carry = 0;
    __syncthreads();
    for(uint32_t base = 0; base < count; base += 1024)
    {
        uint32_t i = base + threadIdx.x;
        uint32_t v = (i < count) ? data[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for(uint32_t off = 1; off < 1024; off <<= 1)
        {
            uint32_t add = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        uint32_t incl = s[threadIdx.x];
        if(i < count) data[i] = carry + incl - v;
        __syncthreads();
        if(threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if(total && threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kRsThreads)
KCRadixScatter(const uint64_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
               uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, uint32_t n,
               uint32_t shift, uint32_t numTiles, const uint32_t* __restrict__ histScan)
{
    __shared__ uint32_t base[256];
    __shared__ uint32_t waveCnt[kRsWaves][256];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint64_t ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    base[tid] = histScan[size_t(tid) * numTiles + blockIdx.x];
    uint32_t running = 0;   // keys of digit `tid` already placed by this tile
    const uint32_t tileStart = blockIdx.x * kRsTile;
    for(uint32_t r = 0; r < kRsRounds; r++)
    {
        for(uint32_t w = 0; w < kRsWaves; w++) waveCnt[w][tid] = 0;
        __syncthreads();
        uint32_t i = tileStart + r * kRsThreads + tid;
        bool valid = (i < n);
        uint64_t key = valid ? keysIn[i] : 0ull;
        uint32_t val = valid ? valsIn[i] : 0u;
        uint32_t digit = uint32_t(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        #pragma unroll
        for(uint32_t b = 0; b < 8; b++)
        {
            uint64_t bb = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bb : ~bb;
        }
        uint32_t rank = uint32_t(__popcll(peers & ltMask));
        if(valid && rank == 0) waveCnt[wave][digit] = uint32_t(__popcll(peers));
        __syncthreads();
        {
            uint32_t acc = running;
            for(uint32_t w = 0; w < kRsWaves; w++)
            {
                uint32_t c = waveCnt[w][tid];
                waveCnt[w][tid] = acc;
                acc += c;
            }
            running = acc;
        }
        __syncthreads();
        if(valid)
        {
            uint32_t dst = base[digit] + waveCnt[wave][digit] + rank;
            keysOut[dst] = key;
            valsOut[dst] = val;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// KCConstructLBVHInternalNodes (AcceleratorLBVH.cu:170-299), Karras 2012
// Fig. 4 with the reference's Delta (equal codes fall back to the index XOR,
// AcceleratorLBVH.cu:69-94).  Also records each node's leaf range.
// ---------------------------------------------------------------------------
// strict = 0: the reference's fallback, clz(i ^ j) on equal codes.  It can
// rank a run of equal codes below neighbours that differ only in low bits,
// which breaks the Karras prefix order (some nodes/leaves get no parent).
// strict = 1: the key is (code, index), so equal codes rank above every
// distinct pair (64 + clz32(i ^ j)); used to rebuild such inputs.
__device__ __forceinline__ int32_t Delta(const uint64_t* __restrict__ codes, int32_t n,
                                         int32_t i, int32_t j, uint32_t strict)
{
    if(j < 0 || j >= n) return -1;
    uint64_t l = codes[i], r = codes[j];
    if(l == r)
    {
        if(strict) return 64 + int32_t(__clz(uint32_t(i ^ j)));
        l = uint64_t(i);
        r = uint64_t(j);
    }
    uint64_t d = l ^ r;
    return (d == 0) ? 64 : int32_t(__clzll(d));
}

__device__ __forceinline__ int32_t DivUpI(int32_t v, int32_t d) { return (v + d - 1) / d; }

__global__ void __launch_bounds__(kTPB)
KCKarras(const uint64_t* __restrict__ codes, const uint32_t* __restrict__ leafOrder, int32_t n,
         uint32_t* __restrict__ nodes, uint32_t* __restrict__ leafParent,
         uint32_t* __restrict__ nodeRange, uint32_t strict)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if(n == 1)
    {
        if(i == 0)
        {
            nodes[0] = kLeafFlag | 0u;
            nodes[1] = kInvalid;
            nodes[2] = kInvalid;
            leafParent[leafOrder[0]] = 0;
            if(nodeRange) { nodeRange[0] = 0; nodeRange[1] = 0; }
        }
        return;
    }
    if(i >= n - 1) return;
    int32_t diff = Delta(codes, n, i, i + 1, strict) - Delta(codes, n, i, i - 1, strict);
    int32_t d = (diff < 0) ? -1 : 1;
    int32_t deltaMin = Delta(codes, n, i, i - d, strict);
    int32_t lMax = 2;
    while(Delta(codes, n, i, i + lMax * d, strict) > deltaMin) lMax <<= 1;
    int32_t l = 0;
    for(int32_t t = lMax >> 1; t != 0; t >>= 1)
        if(Delta(codes, n, i, i + (l + t) * d, strict) > deltaMin) l += t;
    int32_t j = i + l * d;
    int32_t s = 0;
    int32_t deltaNode = Delta(codes, n, i, j, strict);
    for(int32_t t = DivUpI(l, 2); t != 0; t = (t == 1) ? 0 : DivUpI(t, 2))
        if(Delta(codes, n, i, i + (s + t) * d, strict) > deltaNode) s += t;
    int32_t gamma = i + s * d + min(d, 0);

    uint32_t* me = nodes + size_t(i) * 3;
    int32_t lo = min(i, j), hi = max(i, j);
    if(lo == gamma)
    {
        uint32_t leaf = leafOrder[gamma];
        me[0] = kLeafFlag | leaf;
        leafParent[leaf] = uint32_t(i);
    }
    else
    {
        me[0] = uint32_t(gamma);
        nodes[size_t(gamma) * 3 + 2] = uint32_t(i);
    }
    if(hi == gamma + 1)
    {
        uint32_t leaf = leafOrder[gamma + 1];
        me[1] = kLeafFlag | leaf;
        leafParent[leaf] = uint32_t(i);
    }
    else
    {
        me[1] = uint32_t(gamma + 1);
        nodes[size_t(gamma + 1) * 3 + 2] = uint32_t(i);
    }
    if(i == 0) me[2] = kInvalid;
    if(nodeRange)
    {
        nodeRange[2 * i + 0] = uint32_t(lo);
        nodeRange[2 * i + 1] = uint32_t(hi);
    }
}

// ---------------------------------------------------------------------------
// KCUnionLBVHBoundingBoxes (AcceleratorLBVH.cu:301-435): one thread per leaf
// climbs; the second arrival at a node (acq_rel agent-scope counter) unions
// both children and continues.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kTPB)
KCRefit(const uint32_t* __restrict__ nodes, const uint32_t* __restrict__ leafParent,
        const float* __restrict__ leafAabbs, uint32_t n, uint32_t* __restrict__ arrive,
        float* __restrict__ boxes)
{
    uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if(leaf >= n) return;
    const uint32_t nodeCount = (n > 1) ? n - 1 : 1;
    uint32_t node = leafParent[leaf];
    // A well-formed tree is at most n levels deep.  The step bound and the
    // index checks keep a malformed topology (see Delta) in bounds and finite;
    // the host validates the tree afterwards.
    for(uint32_t steps = 0; node < nodeCount && steps <= n; steps++)
    {
        uint32_t old = __hip_atomic_fetch_add(&arrive[node], 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t* nd = nodes + size_t(node) * 3;
        bool single = (nd[1] == kInvalid);
        if(old == 0 && !single) return;
        if(old > 1) return;
        float b[6];
        for(int c = 0; c < (single ? 1 : 2); c++)
        {
            uint32_t ch = nd[c];
            bool ok = (ch & kLeafFlag) ? ((ch & ~kLeafFlag) < n) : (ch < nodeCount);
            if(!ok) return;
            const float* src = (ch & kLeafFlag) ? (leafAabbs + 6 * size_t(ch & ~kLeafFlag))
                                                : (boxes + 6 * size_t(ch));
            float cb[6];
            for(int k = 0; k < 6; k++)
                cb[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if(c == 0)
                for(int k = 0; k < 6; k++) b[k] = cb[k];
            else
            {
                for(int k = 0; k < 3; k++)
                {
                    b[k] = MinF(b[k], cb[k]);
                    b[3 + k] = MaxF(b[3 + k], cb[3 + k]);
                }
            }
        }
        float* dst = boxes + 6 * size_t(node);
        for(int k = 0; k < 6; k++)
            __hip_atomic_store(dst + k, b[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        node = nd[2];
    }
}

__global__ void __launch_bounds__(kTPB) KCIota(uint32_t* __restrict__ out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) out[i] = i;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
int LaunchIota(uint32_t* dOut, uint32_t n, hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    KCIota<<<DivUp(n, kTPB), kTPB, 0, s>>>(dOut, n);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

int LaunchPrimCentersAabbs(const float* dPos, const uint32_t* dIdx, uint32_t n, float* dCenters,
                           float* dAabbs, hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    KCPrimCentersAabbs<<<DivUp(n, kTPB), kTPB, 0, s>>>(dPos, dIdx, n, dCenters, dAabbs);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

int LaunchReduceAabb(const float* dAabbs, uint32_t n, float* dOut, float* dPartial, hipStream_t s)
{
    uint32_t blocks = std::min<uint32_t>(kReduceBlocks, std::max<uint32_t>(1u, DivUp(n, kTPB)));
    KCReduceAabbPartial<<<blocks, kTPB, 0, s>>>(dAabbs, n, dPartial);
    MRAY_HIP_CHECK(hipGetLastError());
    KCReduceAabbPartial<<<1, kTPB, 0, s>>>(dPartial, blocks, dOut);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

size_t ReduceAabbTempBytes() { return size_t(kReduceBlocks) * 6 * sizeof(float); }

int LaunchMorton63(const float* dCenters, uint32_t n, const float* dBox, uint64_t* dCodes,
                   hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    KCMorton63<<<DivUp(n, kTPB), kTPB, 0, s>>>(dCenters, n, dBox, dCodes);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

int LaunchMorton30(const float* dCenters, uint32_t n, const float* dBox, uint32_t* dCodes,
                   hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    KCMorton30<<<DivUp(n, kTPB), kTPB, 0, s>>>(dCenters, n, dBox, dCodes);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

size_t RadixSortTempBytes(uint32_t n)
{
    uint32_t tiles = std::max(1u, DivUp(n, kRsTile));
    size_t hist = size_t(256) * tiles * sizeof(uint32_t);
    size_t kv = size_t(n) * (sizeof(uint64_t) + sizeof(uint32_t));
    return ((hist + 255) / 256) * 256 + kv + 256;
}

int LaunchRadixSort(uint64_t* dKeys, uint32_t* dVals, uint32_t n, void* dTemp, size_t tempBytes,
                    hipStream_t s)
{
    if(n <= 1) return MRAY_OK;
    if(tempBytes < RadixSortTempBytes(n))
    {
        SetError("radix sort: temporary buffer too small");
        return MRAY_ERR_INVALID;
    }
    uint32_t tiles = DivUp(n, kRsTile);
    char* t = static_cast<char*>(dTemp);
    uint32_t* hist = reinterpret_cast<uint32_t*>(t);
    size_t histBytes = ((size_t(256) * tiles * sizeof(uint32_t) + 255) / 256) * 256;
    uint64_t* altK = reinterpret_cast<uint64_t*>(t + histBytes);
    uint32_t* altV = reinterpret_cast<uint32_t*>(altK + n);
    uint64_t* srcK = dKeys;
    uint32_t* srcV = dVals;
    uint64_t* dstK = altK;
    uint32_t* dstV = altV;
    for(uint32_t pass = 0; pass < 8; pass++)
    {
        uint32_t shift = pass * 8;
        KCRadixHist<<<tiles, kRsThreads, 0, s>>>(srcK, n, shift, tiles, hist);
        MRAY_HIP_CHECK(hipGetLastError());
        KCExclusiveScanSingle<<<1, 1024, 0, s>>>(hist, 256 * tiles, nullptr);
        MRAY_HIP_CHECK(hipGetLastError());
        KCRadixScatter<<<tiles, kRsThreads, 0, s>>>(srcK, srcV, dstK, dstV, n, shift, tiles, hist);
        MRAY_HIP_CHECK(hipGetLastError());
        std::swap(srcK, dstK);
        std::swap(srcV, dstV);
    }
    // 8 passes: the result is back in the caller's buffers
    return MRAY_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free sort (mray_radix_sort_u64_u32_noalloc): the same passes over
// a scratch area that lives in the code object, one copy per GPU, held until
// the process ends.  Its size is RadixSortTempBytes(kNoAllocSortMax).
// ---------------------------------------------------------------------------
constexpr uint32_t kNoAllocSortMax = MRAY_NOALLOC_SORT_MAX;
static_assert(kNoAllocSortMax == (1u << 20), "scratch size is stated as 2^20 pairs");
constexpr size_t kNoAllocSortHistBytes =
    ((size_t(256) * ((kNoAllocSortMax + kRsTile - 1) / kRsTile) * sizeof(uint32_t) + 255) / 256) * 256;
constexpr size_t kNoAllocSortBytes =
    kNoAllocSortHistBytes + size_t(kNoAllocSortMax) * (sizeof(uint64_t) + sizeof(uint32_t)) + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocSortBytes];

namespace
{
constexpr int kMaxDevices = 64;
std::mutex gScratchLock;
void* gScratchPtr[kMaxDevices] = {};
}

// The device address of `scratch` on the current GPU, looked up once per GPU
// (the array's name is only a host-side handle; kernels need this address).
static int ScratchOnCurrentDevice(void** out)
{
    int dev = 0;
    MRAY_HIP_CHECK(hipGetDevice(&dev));
    if(dev < 0 || dev >= kMaxDevices)
    {
        SetError("radix sort (no-alloc): device index out of range");
        return MRAY_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(gScratchLock);
    if(!gScratchPtr[dev])
    {
        void* p = nullptr;
        MRAY_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if(!p)
        {
            SetError("radix sort (no-alloc): scratch symbol has no device address");
            return MRAY_ERR_HIP;
        }
        gScratchPtr[dev] = p;
    }
    *out = gScratchPtr[dev];
    return MRAY_OK;
}

size_t LbvhTempBytes(uint32_t n)
{
    // leaf parents [n] + arrival counters [n]
    return size_t(std::max(n, 1u)) * 2 * sizeof(uint32_t) + 256;
}

int LaunchLbvhBuild(const uint64_t* dSortedCodes, const uint32_t* dLeafOrder,
                    const float* dLeafAabbs, uint32_t n, uint32_t* dNodes, float* dBoxes,
                    uint32_t* dNodeRange, void* dTemp, size_t tempBytes, hipStream_t s,
                    uint32_t strictTie)
{
    if(n == 0) return MRAY_OK;
    if(tempBytes < LbvhTempBytes(n))
    {
        SetError("lbvh build: temporary buffer too small");
        return MRAY_ERR_INVALID;
    }
    uint32_t nodeCount = (n > 1) ? n - 1 : 1;
    uint32_t* leafParent = static_cast<uint32_t*>(dTemp);
    uint32_t* arrive = leafParent + std::max(n, 1u);
    MRAY_HIP_CHECK(hipMemsetAsync(dNodes, 0xFF, size_t(nodeCount) * 3 * sizeof(uint32_t), s));
    MRAY_HIP_CHECK(hipMemsetAsync(arrive, 0, size_t(nodeCount) * sizeof(uint32_t), s));
    MRAY_HIP_CHECK(hipMemsetAsync(leafParent, 0xFF, size_t(n) * sizeof(uint32_t), s));
    KCKarras<<<DivUp(std::max(n - 1, 1u), kTPB), kTPB, 0, s>>>(dSortedCodes, dLeafOrder,
                                                               int32_t(n), dNodes, leafParent,
                                                               dNodeRange, strictTie);
    MRAY_HIP_CHECK(hipGetLastError());
    KCRefit<<<DivUp(n, kTPB), kTPB, 0, s>>>(dNodes, leafParent, dLeafAabbs, n, arrive, dBoxes);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

} // namespace mray

using namespace mray;

extern "C" int mray_radix_sort_u64_u32_noalloc(uint64_t* dKeys, uint32_t* dValues, uint32_t count,
                                                void* stream)
{
    if(count > kNoAllocSortMax)
    {
        SetError("radix sort (no-alloc): " + std::to_string(count) + " pairs exceed the maximum of " +
                 std::to_string(kNoAllocSortMax));
        return MRAY_ERR_INVALID;
    }
    if(count > 1 && (!dKeys || !dValues))
    {
        SetError("radix sort (no-alloc): NULL keys or values");
        return MRAY_ERR_INVALID;
    }
    if(count <= 1) return MRAY_OK;
    void* tmp = nullptr;
    MRAY_TRY(ScratchOnCurrentDevice(&tmp));
    return LaunchRadixSort(dKeys, dValues, count, tmp, kNoAllocSortBytes,
                           static_cast<hipStream_t>(stream));
}

