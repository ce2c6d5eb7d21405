// dev_types.h — HBM layout of the scene and of the wavefront queues, shared
// by the host API (which fills it) and the kernels (which read it).
//
// Scene (read-only, built once by upload_scene):
//   bnodes  2 x float4 per BLAS node, all meshes concatenated:
//           {lo.x, lo.y, lo.z, hi.x} {hi.y, hi.z, child, entry}
//           interior: child = global index of the left child (right = +1);
//           leaf: child = global triangle slot of its first triangle;
//           entry = the node's traversal stack entry (see dtrace.h)
//   bigcnt  per global triangle slot: the triangle count of a leaf of 15
//           or more triangles starting there (0 elsewhere)
//           Each mesh starts at an odd global index, so every child pair
//           (left index odd within the mesh) is 64-byte aligned.
//   tris    3 x float4 per leaf slot, in leaf order:
//           {v0.xyz, e1.x} {e1.yz, e2.xy} {e2.z, tri, 0, 0}
//           (e1 = v1 - v0, e2 = v2 - v0, computed once on the host: the same
//           IEEE subtraction the reference performs per test)
//   tnodes  3 x float4 per TLAS node: the two BLAS-style float4 plus
//           {centre of left child, centre of right child, axis, 0} along the
//           node's longest axis, the traversal order key of tlas.cpp:231-241
//   inst    4 x float4 per instance: world_to_local columns c0..c3 packed
//           {c0.xyz, c1.x} {c1.yz, c2.xy} {c2.z, c3.xyz} and
//           {blas root entry, instance_id, tri_base | flat<<31, light map or -1}
//           (the root's stack entry, see kEntShift; kNoRoot for an empty mesh)
//   vidx    3 x int per global triangle (global vertex indices)
//   vnrm    float4 per global vertex (xyz normal)
//   tmat    material id per global triangle
//   lmap    per instance whose id names area lights: first light index per
//           mesh triangle, -1 for none
//   lights  4 x float4 per light: {p0, area} {p1, e.r} {p2, e.g} {n_geo, e.b}
//   mats    4 x float4 per material (see DevMat)
//   suns    2 x float4 per directional light: {dir, 0} {emission, 0}
//   prims   analytic primitives with their PrimBLAS (pnodes, porder); the
//           instance of the primitives has root kPrimRoot
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace xnh {

constexpr int kMaxStack = 128;          // TLAS 64 + BLAS 64, the reference's two stacks
constexpr uint32_t kFlatBit = 0x80000000u;
// BLAS stack entries: count << kEntShift | index (see dtrace.h)
constexpr uint32_t kEntShift = 28;
constexpr uint32_t kEntIndexMask = (1u << kEntShift) - 1u;
constexpr uint32_t kEntBigLeaf = 15;
constexpr uint32_t kNoRoot = 0xFFFFFFFFu;
// Instance root of the scene's analytic primitives (its PrimBLAS), and the
// flag that marks a primitive in a hit's slot (slot = primitive index).
constexpr uint32_t kPrimRoot = 0xFFFFFFFEu;
constexpr uint32_t kPrimHit = 0x80000000u;
// Bound on the traversal iterations of one wave in a launch: far above any
// legitimate walk (a wave's share of a 2^28-path batch of the 4096-instance
// scene takes some 3e5 iterations), it makes every wave exit even if its
// walks stopped making progress; a wave that reaches it abandons its work
// and is reported. (Every iteration advances at least one lane by one step,
// and the host rejects trees whose children do not follow their parent.)
constexpr uint32_t kWaveStepLimit = 1u << 26;

struct DevSky {
  float zenith[3], horizon[3], ground[3];
  float sharpness, intensity;
};

struct DevScene {
  const float4* bnodes;
  const uint32_t* bigcnt;
  const float4* tris;
  const float4* tnodes;
  const uint32_t* torder;
  const float4* inst;
  const int32_t* vidx;
  const float4* vnrm;
  const int32_t* tmat;
  const int32_t* lmap;
  const float4* lights;
  const float4* mats;
  const float4* suns;
  const float4* prims;      // 3 x float4 per primitive: {type, mat, p0, p1} {p2..p5} {p6..p9}
  const float4* pnodes;     // 2 x float4 per PrimBLAS node, as bnodes without the pair layout
  const uint32_t* porder;   // PrimBLAS leaf slot -> primitive index
  const uint32_t* ikind;    // per instance: the shading class of all its surfaces (kShade*), for sorting
  int32_t num_pnodes;
  int32_t num_tnodes;
  int32_t num_lights;
  int32_t num_suns;
  int32_t pad;
  DevSky sky;
};

// Shading classes: which branch of the shading kernel a hit takes, so that
// a wave can be given hits of one class (kernels.hip k_shade). An instance's
// class is its materials' common class, or kShadeMixed.
enum ShadeClass : uint32_t {
  kShadeMiss = 0,
  kShadeLambert,      // opaque, non-metal, F0 = 0: one Lambert lobe
  kShadeConductor,
  kShadeDielectric,   // opaque dielectric: GGX and Lambert lobes
  kShadeRoughGlass,
  kShadeDelta,
  kShadeSubsurface,
  kShadeMixed,        // mixed classes, or analytic primitives
  kShadeInvalid,      // no entry (past the end of the queue)
  kNumShadeClasses
};

// float4 per material record (device/dbsdf.h DMat)
constexpr int kMatVecs = 6;

struct DevCamera {
  float eye[3], llc[3], horiz[3], vert[3];
};

// Wavefront state. A render call is a list of jobs (scene, camera,
// resolution, sample range, depth); the host cuts their samples into
// batches of chunks, each chunk a contiguous sample range of one job, and
// uploads all chunk tables once. Chunk c of a batch owns the per-sample
// slots [base, base + cap) (cap = pixels * samples; slot order is
// sample-major: slot = base + s * pixels + pixel) and, in both path queues,
// kQSeg segments of qseg entries from qbase (in the shadow queue twice
// that, from 2 qbase): per-job queue segments, so every 64-entry unit of
// work belongs to exactly one job and a wave reads its job's scene through
// scalar loads. The shading kernel appends into the segment of its
// work-group's group (block index mod kQSeg), each segment with its own
// counter kCtrStride words from the next: one counter for all appends of a
// chunk served too few atomics per second (one address is resolved past the
// XCDs' L2 caches, one operation at a time). Ray generation writes depth 0
// as segment 0, in slot order.
constexpr int kMaxChunks = 64;
constexpr int kQSeg = 8;
constexpr int kCtrStride = 64;   // words between two segment counters (256 B)
// Entries of a segment beyond cap / kQSeg. The shading kernel's grid holds a
// multiple of 32 waves, and window k (eight 64-entry units) goes to wave
// k mod waves, whose group is (wave / 4) mod kQSeg: of n consecutive
// windows a group takes at most n / 8 + 4, so its outputs for a chunk of
// cnt <= cap input entries (at most one continuation, two shadow rays per
// entry) stay below cap / 8 + 64 + 4.25 * 512 per segment.
constexpr int kSegSlack = 2304;
// Segment size of a chunk of cap paths (a multiple of 64).
constexpr int64_t qseg_for(int64_t cap) { return ((cap + kQSeg - 1) / kQSeg + kSegSlack + 63) / 64 * 64; }
// Queue entries beyond the slots that a batch of up to kMaxChunks chunks
// needs for its segments.
constexpr int64_t kQueueSlack = int64_t(kMaxChunks) * kQSeg * (kSegSlack + 128);

// Division of n < 2^31 by a divisor d that is fixed for a batch, as one
// 32x32->64-bit multiply and a shift: n / d = (n * m) >> p with
// p = 31 + ceil(log2 d) and m = ceil(2^p / d) (< 2^32). m exceeds 2^p / d
// by less than 1, so n * m / 2^p exceeds n / d by less than n / 2^p < 1/d
// and the floor is exact (division by invariant integers using
// multiplication, Granlund and Montgomery 1994).
struct FastDiv {
  uint32_t m, p;
};
inline FastDiv fast_div(uint32_t d) {
  const uint32_t k = d <= 1u ? 0u : 32u - static_cast<uint32_t>(__builtin_clz(d - 1u));
  const uint32_t p = 31u + k;
  return FastDiv{static_cast<uint32_t>(((uint64_t(1) << p) + d - 1u) / d), p};
}

struct DevJob {
  DevScene S;
  DevCamera cam;
  int32_t W, H, minb, maxb;
  float4* accum;    // per-pixel sums of this job
  FastDiv band;     // division by 8 W (the rows of pixel tiles)
};

struct DevChunk {
  int32_t job, spp_first, nsamples, npix;
  int32_t base, cap;
  int32_t qbase, qseg;   // the chunk's path-queue segments: kQSeg of qseg entries from qbase
  int32_t cam_lg;        // depth-0 entry order: log2 of the samples per 64-entry block (3: 8 samples of
                         // 8 pixels, 6: 64 samples of one pixel), 0 = sample-major
  FastDiv sdiv;          // division by npix (cam_lg 0) or by the pixel blocks npix >> (6 - cam_lg)
};

// Bounce paths only: depth-0 (camera) paths are rebuilt from their entry
// index and store nothing here.
struct PathQueue {
  float4* ray0;     // {o.xyz, d.x}
  float4* ray1;     // {d.y, d.z, prev_pdf, flags: bit 0 delta, bits 1-3 direction octant}
  float4* thr;      // {T.xyz, 0}
  uint4* st;        // {rng lo, rng hi, slot, training record or kNoRecord}
};

struct WaveBuffers {
  PathQueue q[2];
  float4* hit;      // {t, u, v, instance index (bits)} per queue entry; -1 = miss
  uint32_t* hit_slot;  // global triangle slot per queue entry
  float4* sh0;      // shadow ray {o.xyz, d.x}
  float4* sh1;      // {d.y, d.z, tmax, slot*2+k}
  float4* sh_c;     // unoccluded contribution
  float4* l_term;   // per slot: emission / sky at path end
  float4* l_nee;    // per slot x 2: area and sun NEE sums
  const DevJob* jobs;
  const DevChunk* chunks;   // [batch][kMaxChunks]
  const int32_t* nchunks;   // chunks per batch
  uint32_t* counters;       // see Ctr
  unsigned long long* stats;  // see Stat
  uint32_t* spill;  // traversal stack entries beyond the LDS part
  int32_t capacity;   // slots (paths) of a batch
  int32_t qcap;       // entries of each path-queue array (capacity + kQueueSlack); the shadow queue holds twice
  // path guiding (null / 0 when off)
  float4* mix;      // [kMixVecs][mix_stride] mixture per queue entry (k_guide_infer)
  uint32_t mix_stride;  // queue entries a guided batch uses at most (<= capacity): the stride of mix and of
                        // the hash-grid encodings
  float4* rec;      // [4][rec_cap] training records, see guide.h
  uint32_t rec_cap;  // a multiple of kQSeg
  uint32_t rec_every;   // a path records its vertices when hash(slot, batch) % rec_every == 0
  float guide_alpha;    // largest one-sample MIS probability of the guide; 0 = guiding off
  const float* alpha_now;   // the probability used by the current batch (device memory)
  int32_t guide_depth;      // bounces below this depth are guided
};

enum Ctr : int {
  kBatch = 0,                            // batch being rendered (advanced on the device)
  kOverflow = 1,                         // lanes whose stack exceeded kMaxStack
  kWatchdog = 3,                         // waves that hit an iteration bound (see dtrace.h)
  kSegOverflow = 4,                      // appends past a queue segment's end (dropped; never expected)
  kCnt = 16,                                        // [queue][chunk][segment] live paths (stride kCtrStride)
  kShCnt = kCnt + 2 * kMaxChunks * kQSeg * kCtrStride,  // [chunk][segment] shadow rays
  kCursor = kShCnt + kMaxChunks * kQSeg * kCtrStride,   // per-launch work cursors of a batch
  kNumCursors = 2 * 64 + 4,
  // [segment] guiding training records of the batch (stride kCtrStride)
  kRecCnt = (kCursor + kNumCursors + kCtrStride - 1) / kCtrStride * kCtrStride,
  kNumCtr = kRecCnt + kQSeg * kCtrStride
};

// Counter of segment x of chunk ci: path queue q, shadow queue.
constexpr int ctr_path(int q, int ci, int x) { return kCnt + ((q * kMaxChunks + ci) * kQSeg + x) * kCtrStride; }
constexpr int ctr_shadow(int ci, int x) { return kShCnt + (ci * kQSeg + x) * kCtrStride; }
// 64-entry units per window of the shading kernel
constexpr uint32_t kShadeWin = 8;

// Counter of training-record segment x; the records are kQSeg segments of
// rec_cap / kQSeg, segment x from x * (rec_cap / kQSeg). A shading window
// appends its records to segment (window index) % kQSeg.
constexpr int ctr_rec(int x) { return kRecCnt + x * kCtrStride; }

enum Stat : int {
  kStPaths = 0,
  kStClosest,
  kStShadow,
  kStNodeTests,
  kStInterior,
  kStTriTests,
  kStInstVisits,
  kStShNodeTests,     // the same four for any-hit (shadow) traversal
  kStShInterior,
  kStShTriTests,
  kStShInstVisits,
  kStNonFinite,
  // SIMD use of the traversal (counting variants only): wave iterations of
  // the interior and leaf phases and of the persistent round loop (counted
  // once per wave), and the lane-level leaf steps and rounds; closest hit,
  // then any hit
  kStIntIters,
  kStLeafIters,
  kStLeafSteps,
  kStRoundIters,
  kStRoundLanes,
  kStShIntIters,
  kStShLeafIters,
  kStShLeafSteps,
  kStShRoundIters,
  kStShRoundLanes,
  kNumStat
};

// Output stage (image.hip): the device state of one NoiseStats computation
// (math/image.h:11-31 plus the radix-select state) and its scratch.
constexpr int kImgMaxBlocks = 2048;                 // partial-sum blocks of one pass
constexpr int64_t kImgMaxPixels = int64_t(1) << 26; // 8192 x 8192
struct ImgState {
  float avg, edge, eps;        // average luminance, edge threshold, epsilon
  float max_abs, mean_sq, mean_abs;
  uint32_t count, pad;         // counted (interior, non-edge) pixels
  uint32_t rank[2];            // median, 95th percentile: rank left inside the prefix
  uint32_t prefix[2];          // bits of the selected key found so far
  uint32_t hist[512];          // [target][digit] of the current pass
};
// The allocation-free NoiseStats entry point's static scratch: images of at
// most 2^22 pixels (2048 x 2048), luminance and keys 16 MiB each plus the
// per-block partials and the state, 256-byte aligned.
constexpr int64_t kImgStaticMaxPixels = int64_t(1) << 22;
constexpr size_t kImgStaticScratchBytes = 33605888;   // 32.05 MiB
struct ImgScratch {
  float* lum;                  // W*H luminance
  uint32_t* key;               // W*H residual keys
  double* part;                // 2 * kImgMaxBlocks partial sums
  float* pmax;                 // kImgMaxBlocks partial maxima
  uint32_t* pcnt;              // kImgMaxBlocks partial counts
  ImgState* st;
};

}  // namespace xnh
