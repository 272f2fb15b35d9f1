// Host-side engine objects behind the C-ABI: device buffers, the device scene
// (reference: CUDA::DeviceSceneDescription, DeviceSceneDescription.h:13-21),
// the LBVH builder (CUDA::BVHBuilder, BVHBuilder.h:12-19) and the progressive
// tracer (CUDAPathTracer, cuda_path_tracer.h:13-40).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <map>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

#include "pt_api.h"

namespace pt {

// Errors inside the library are C++ exceptions; the C-ABI layer converts them
// to status codes and never lets one cross the ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PT_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::pt::Error(PT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Grow-only device allocation.
class DevBuf {
   public:
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void reserve(size_t bytes) {
        if (bytes <= cap_) return;
        release();
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            throw Error(PT_E_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
        ptr_ = p;
        cap_ = bytes;
    }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        cap_ = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(ptr_); }
    size_t capacity() const { return cap_; }
    friend void swap(DevBuf& a, DevBuf& b) noexcept {
        std::swap(a.ptr_, b.ptr_);
        std::swap(a.cap_, b.cap_);
    }

   private:
    void* ptr_ = nullptr;
    size_t cap_ = 0;
};

// Pinned host staging (hipHostMalloc), grow-only.
class HostBuf {
   public:
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() {
        if (ptr_) (void)hipHostFree(ptr_);
    }
    void reserve(size_t bytes) {
        if (bytes <= cap_) return;
        if (ptr_) (void)hipHostFree(ptr_);
        ptr_ = nullptr;
        cap_ = 0;
        PT_HIP(hipHostMalloc(&ptr_, bytes, hipHostMallocDefault));
        cap_ = bytes;
    }
    template <class T>
    T* as() const { return static_cast<T*>(ptr_); }

   private:
    void* ptr_ = nullptr;
    size_t cap_ = 0;
};

// LBVH in device memory.  Two layouts live side by side: the reference's
// parity arrays (keys, perm, children, parents) and the traversal layout:
//   nodes[4*i .. 4*i+3]   one 64-byte record per internal node i:
//       (L.minx, L.maxx, L.miny, L.maxy) (L.minz, L.maxz, R.minx, R.maxx)
//       (R.miny, R.maxy, R.minz, R.maxz) (children word, 0, 0, 0)
//     -- both child boxes next to the child word, so one visit is one
//        contiguous 64-byte read; a leaf child's slot holds its triangle box.
//   tris[3*k .. 3*k+2]   one 48-byte record per sorted triangle k:
//       (p0.xyz, original id) (e1.xyz, 0) (e2.xyz, 0), e = p - p0.
struct DeviceBvh {
    uint32_t n = 0;
    DevBuf keys, vals, keys_alt, vals_alt;  // sorted Morton keys, perm (+ ping-pong)
    DevBuf scene_of;                        // scene of each leaf slot
    DevBuf children, parents_leaf, parents_internal, counters;
    DevBuf nodes, tris, root_box;
    DevBuf tile_hist, digit_total;  // radix sort scratch
    DevBuf span_list, span_top, span_count;  // refit: per-level node lists
    DevBuf mid_links, top_links;             // refit: the level-1 and level-2 lists' staged links
    DevBuf slots;                            // refit: each listed node's position in its list
};

// One or more scenes: scene s owns triangles [off[s], off[s+1]) and its own
// LBVH over them (internal node off[s] is its root).
struct DeviceScene {
    uint32_t n = 0;
    uint32_t ninfos = 0;
    std::vector<uint32_t> off_host{0, 0};  // nscenes + 1 offsets
    std::vector<float> box_host;           // per scene: min.xyz, max.xyz
    DevBuf off;                            // device copy of off_host
    DevBuf morton_box;                     // per scene: min.xyz, 1/(max-min).xyz
    DevBuf tris;                           // pt_triangle[n] as uploaded
    DevBuf infos;                          // pt_triangle_info[ninfos]
    DevBuf lights;                         // pt_dir_light[nlights] (all scenes)
    uint32_t nlights = 0;
    DevBuf detail;                         // pt_triangle_detail[ndetail], or none
    DevBuf texels;                         // RGBA8 texels of all textures
    DevBuf texinfo;                        // uint4 per texture: first texel, w, h, 0
    uint32_t ndetail = 0, ntex = 0;
    uint32_t max_id = 0;                   // largest triangle id of the upload
    // Status word the kernels OR PT_STATUS_* flags into (a violated bound:
    // traversal stack, refit list); sticky until pt_get_status reads it.
    DevBuf status;
    DeviceBvh bvh;
    bool ready = false;
    uint32_t nscenes() const { return static_cast<uint32_t>(off_host.size()) - 1; }
};

// Builder timings (ms) of the last build, from HIP events.
struct BuildTimes {
    float total = 0.0f;
};

// LBVH build: Morton keys -> 4-pass LDS radix sort -> radix tree -> refit.
void build_lbvh(DeviceScene& scene, hipStream_t stream);
// Parity dump of one scene in local indices (device -> host, synchronous).
void dump_bvh(DeviceScene& scene, uint32_t scene_index, hipStream_t stream, uint32_t* morton,
              uint32_t* perm, uint32_t* children, uint32_t* parents_leaf,
              uint32_t* parents_internal, float* bboxes);

// LBVH build options (tunable).
struct BuildTuning {
    uint32_t staged = 1;  // refit the upper levels with their node lists staged in LDS (0: in global memory)
};
BuildTuning& build_tuning();

// Launch shape of the persistent closest-hit kernel (tunable).
struct TraceTuning {
    uint32_t blocks = 2048;  // persistent workgroups of 256 threads
    uint32_t refill = 32;    // refill idle lanes when fewer than this many are busy
    uint32_t lds = 16;       // traversal stack entries kept in LDS: 8, 16 or 32
    uint32_t pool = 256;     // rays a wave takes per hand-out on long queues (64..4096)
    uint32_t filter = 1;     // closest-hit batches: root filter before the persistent loop
};
TraceTuning& trace_tuning();

// Bytes of one persistent-traversal ray queue block (slice counters and
// used-up mask); a launch zeroes its block first, so concurrent launches
// need distinct blocks.
size_t trace_queue_bytes();
// The allocation-free entry point's queue block: trace_scratch_setup (with
// `device` current, once per device, at context creation) resolves the
// device's copy of the library's static scratch array; trace_scratch returns
// it (null before set-up).  Calls that use it overlap on one GPU at their
// peril: they share the one block.
void trace_scratch_setup(int device);
uint32_t* trace_scratch(int device);
// Closest hit over a device ray batch (rays: 6 floats each), rays handed
// out through the queue block `queue`.
void trace_closest(const DeviceScene& scene, uint32_t* queue, const float* d_rays, uint32_t n,
                   uint32_t* d_id, float* d_t, float* d_uv, hipStream_t stream);
// Closest hit with a scene index per ray (batched scenes).
void trace_closest_scenes(const DeviceScene& scene, uint32_t* queue, const float* d_rays,
                          const uint32_t* d_scene, uint32_t n, uint32_t* d_id, float* d_t, float* d_uv,
                          hipStream_t stream);
// Same traversal, also counting inner iterations, box tests and triangle
// tests into d_counters[0..2] (uint64, accumulated) and the deepest
// deferred-node stack into d_counters[3] (maximum).
void trace_closest_counted(const DeviceScene& scene, uint32_t* queue, const float* d_rays, uint32_t n,
                           uint32_t* d_id, float* d_t, float* d_uv,
                           unsigned long long* d_counters, hipStream_t stream);

// Per-pass values of a frame, read by the render kernels from device memory.
struct PassParams {
    float dirs[12];  // corner directions: bottom-left, bottom-right, top-right, top-left
    float eye[3];
    uint32_t pass;   // pass index since the last reset (stratum, RNG key)
    uint32_t seedw;  // 32-bit pass seed (RNG key)
    uint32_t scene;  // scene of a batch traced by this pass
};
constexpr uint32_t kMaxPasses = 1024;  // passes per frame
constexpr int kParamRing = 4;        // pinned PassParams slots in flight

// Progressive renderer state (wavefront queues + FP64 accumulator).
struct Renderer {
    int w = 0, h = 0;
    uint32_t spp = 0;  // passes accumulated since the last reset
    float dirs[12] = {0};
    float eye[3] = {0, 0, 0};
    pt_settings settings{1u, 0xFFFFFFFFu, 3.0f, 1u};
    DevBuf accum;       // double[3*w*h]
    DevBuf queues;      // ray/hit queues and per-path state
    DevBuf counters;    // queue counters
    DevBuf params;      // PassParams[kMaxPasses]
    HostBuf params_host;
    hipEvent_t ring_ev[kParamRing] = {};
    bool ring_used[kParamRing] = {};
    int ring_next = 0;
    void* graph_exec = nullptr;   // hipGraphExec_t of the last captured frame
    alignas(8) char graph_key[256] = {};
    bool timed = false;   // a pass has been bracketed by the context events
    uint32_t last_levels = 0;  // queue levels of the last frame
    uint32_t last_chains = 1;  // level chains of the last frame
    // A plain frame's chains 1.. run on these streams (fork and join events;
    // created before the first frame with that many chains, default priority).
    hipStream_t chain_st[7] = {};
    hipEvent_t chain_fork = nullptr;
    hipEvent_t chain_join[7] = {};
    DevBuf cache_io;           // cache-frame queries and training batch
    // A cache frame traces its training paths, from their primary rays to
    // their ends, on `side` while the pixel paths (two bounce levels, or one
    // path-persistent launch), inference, resolve and accumulation run on
    // the frame's stream (fork and join events); created on the first cache
    // frame, default priority.
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    uint32_t last_train = 0;   // training paths of the last cache frame
    void* cache_graph_exec = nullptr;  // hipGraphExec_t of the last cache frame
    alignas(8) char cache_graph_key[384] = {};
    uint32_t stop_at = 0;      // passes after which render calls stop (0: never)
    double gen_seconds = -1.0; // time to reach stop_at, once reached
    std::chrono::steady_clock::time_point t_start;  // first pass after the last reset
    void* screen = nullptr;    // caller's RGBA32F screen buffer (device), or none
    uint32_t screen_w = 0, screen_h = 0;
    size_t screen_pitch = 0;   // bytes between rows
    Renderer() = default;
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    ~Renderer();
};

struct RenderTuning {
    uint32_t graphs = 1;  // replay each frame's kernels as one captured hipGraph
    uint32_t chains = 4;  // level chains of a plain frame (1..8), one stream each
    uint32_t paths = 1;   // plain frames path-persistent (one launch, no bounce levels)
    uint32_t shade_min = 8;   // path-persistent frames: lanes that shade together (1..64)
};
RenderTuning& render_tuning();

void render_resize(Renderer& r, int w, int h, hipStream_t stream);
void render_reset(Renderer& r, hipStream_t stream);
// npasses 1-spp passes in one frame; dirs = 12 floats per pass (NULL: the
// renderer's corner rays for every pass); seeds[npasses].
void render_passes(Renderer& r, const DeviceScene& scene, uint32_t npasses, const float* dirs,
                   const uint64_t* seeds, const uint32_t* scenes, hipStream_t stream);
// Running-mean image (sum / spp as float, alpha 1) of the accumulator's
// min(w, out_w) x min(h, out_h) corner into a device RGBA32F buffer whose rows
// are `pitch` bytes apart.
void render_display(const Renderer& r, void* out, uint32_t out_w, uint32_t out_h, size_t pitch,
                    hipStream_t stream);
// Rays traced by the last frame (sum of its queue lengths); synchronises.
uint64_t render_last_rays(Renderer& r, hipStream_t stream);

struct RadianceCache;
// One 1-spp cache frame (pixel paths end in the cache at x2; training paths
// extend past x2 and train the cache when PT_CACHE_TRAIN is set).
void render_cache_frame(Renderer& r, const DeviceScene& scene, RadianceCache& cache, uint64_t seed,
                        const pt_cache_settings& cs, hipStream_t stream);
// The first n samples (raw, target) of the last cache frame's training batch.
void render_last_batch(Renderer& r, float* raw, float* target, uint32_t n, hipStream_t stream);
// The last cache frame's pixel queries: raw[12 w h], predictions[3 w h], valid[w h].
void render_last_queries(Renderer& r, float* raw, float* pred, uint32_t* valid, hipStream_t stream);

// Learned radiance cache: the MLP's masters, Adam moments, fp16 fragments
// and training scratch (csrc/mlp.hip).
struct RadianceCache {
    DevBuf master, m, v;   // fp32 parameters and Adam moments
    DevBuf fwd, bwd;       // fp16 weight fragments (forward / backward, 16x16x16)
    DevBuf f32;            // fp16 forward fragments for the 16x16x32 inference
    DevBuf part, loss;     // training scratch: per-workgroup gradient partials, per-sample loss
    DevBuf dstep;          // uint32 Adam step count (device side)
    float box_min[3] = {0, 0, 0}, box_inv[3] = {1, 1, 1};
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
    bool ready = false;
    static int param_count();
    void set_box(const float* bmin, const float* bmax);
    void init(uint64_t seed, hipStream_t st);
    void set_params(const float* host, hipStream_t st);
    void get_params(float* host, hipStream_t st);
    void pack(hipStream_t st, uint32_t advance_step);
    // raw: 12 floats per sample (position, normal, view direction, albedo).
    void infer(const float* d_raw, uint32_t n, float* d_out, hipStream_t st);
    // One Adam step on the batch (d_grad_out == NULL), or only the gradient.
    void train(const float* d_raw, const float* d_target, uint32_t n, float* d_loss,
               float* d_grad_out, hipStream_t st);
    // The same in two halves: scratch for batches up to n (allocates), then
    // launches only (capturable into a graph).
    void prepare_train(uint32_t n);
    void launch_train(const float* d_raw, const float* d_target, uint32_t n, float* d_loss,
                      float* d_grad_out, hipStream_t st);
};

}  // namespace pt

struct pt_context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;
    pt::DeviceScene scene;
    pt::Renderer renderer;
    pt::RadianceCache cache;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Recorded on `stream` after every scene upload and LBVH build: a trace
    // on another stream waits for it first.
    hipEvent_t ready_ev = nullptr;
    // Ray queue block of the closest-hit launches per stream (launches on one
    // stream run in order; on different streams each has its own block).
    std::map<hipStream_t, pt::DevBuf> trace_queues;
    pt::DevBuf io_rays, io_id, io_t, io_uv;  // staging for host-pointer traces
    pt::DevBuf count_buf;                    // counters of pt_trace_closest_counted
    pt_context() = default;
    pt_context(const pt_context&) = delete;
    pt_context& operator=(const pt_context&) = delete;
    // Waits for the context's work, then releases its events and own stream
    // (device buffers are released by their members afterwards).
    ~pt_context() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ready_ev) (void)hipEventDestroy(ready_ev);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};
