// Host runtime of the HIP backend: the C-ABI declared in include/rm_hip.h.
//
// Replaces the reference's Metal host code (src/engine/metal.mm:23-414): one
// context per device holding the resident scene, grow-only frame buffers and
// the temporal history; per frame the same branch decisions (ReSTIR +
// denoise path or megakernel with 1/8/32 lanes), the same pass graph and the
// same history ping-pong, issued on one HIP stream and synchronised once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rm_hip.h"
#include "rm_kernels.h"
#include "rm_kpn.h"
#include "rm_layout.h"
#include "rm_pack.h"
#include "rm_internal.h"

// Work space of rm_trace_device, the entry point that allocates nothing: the
// traversal counters of one call (kNumCounters words, at most kScratchWords =
// 256 bytes). Each GPU holds its own copy until the process ends; every
// rm_trace_device call on a GPU uses that copy, so calls that overlap on one
// GPU race. Host code reaches it only through the address hipGetSymbolAddress
// returns for the device (rm_ctx::scratch_dev), never through the name.
constexpr int kScratchWords = 32;
constexpr size_t kTraceDeviceMaxRays = size_t(1) << 30;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];
static_assert(rmh::kNumCounters <= kScratchWords, "traversal counters fit the scratch array");

namespace {

using namespace rmh;

std::mutex g_create_mu;
std::string g_create_error;
// Device address of `scratch` per HIP device, looked up once per device
// (under g_create_mu) with that device current.
std::vector<void*> g_scratch_dev;

// Grow-only device allocation (the reference's upload/alloc helpers,
// metal.mm:171-178,256-260).
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes) {
        bytes = std::max<size_t>(bytes, 16);
        if (bytes <= cap) return true;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        if (hipMalloc(&ptr, bytes) != hipSuccess) return false;
        cap = bytes;
        // Fresh buffers start zeroed so history that was never written reads
        // as "no data" (kind 0, invalid reservoir, black). hipMemset runs on
        // the null stream, which the context's non-blocking stream does not
        // wait for: finish it here, or it could land after work the caller
        // queues next on that stream (an upload into this buffer).
        return hipMemset(ptr, 0, bytes) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct GBufSet {
    DevBuf p[5];
    bool ensure(size_t px) {
        for (auto& b : p)
            if (!b.ensure(px * sizeof(float4))) return false;
        return true;
    }
    GBufferSoA soa() const {
        return {p[0].as<float4>(), p[1].as<float4>(), p[2].as<float4>(), p[3].as<float4>(), p[4].as<float4>()};
    }
    void release() { for (auto& b : p) b.release(); }
};

struct ResSet {
    DevBuf r[4];
    bool ensure(size_t px) {
        for (auto& b : r)
            if (!b.ensure(px * sizeof(float4))) return false;
        return true;
    }
    ReservoirSoA soa() const { return {r[0].as<float4>(), r[1].as<float4>(), r[2].as<float4>(), r[3].as<float4>()}; }
    void release() { for (auto& b : r) b.release(); }
};

}  // namespace

struct rm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string err;
    bool profiling = false;
    int schedule = 0;  // path passes: 0 persistent lanes, 1 wavefront, 2 wavefront without the persistent tail
    hipEvent_t ev[16] = {};

    // Resident scene.
    bool has_scene = false;
    DevBuf nodes, geo, nrm, nn, matid, mats, uv, lights, cdf, texels;
    DevBuf level_nodes;
    // GPU-built tree: triangle records in scene order, the leaf-order
    // permutation, build scratch and a pinned read-back of level sizes.
    bool gpu_tree = false;
    DevBuf geo_scene, perm, build_scratch;
    // Geometry staged by rm_geo_stage_put for the next GPU-tree upload
    // (device format, scene order): triangles expected, triangles staged, and
    // whether every staging copy was queued.
    DevBuf geo_stage, nrm_stage, nn_stage;
    size_t stage_ntri = 0;
    std::atomic<size_t> staged_tris{0};
    std::atomic<bool> stage_ok{false};
    uint32_t* build_host = nullptr;
    std::vector<int32_t> level_start;  // refit launch ranges into level_nodes, deepest level first
    int32_t nnodes = 0, nlights = 0, opaque = 1, stack_need = 0, ntri = 0;
    float light_total = 0.0f;
    SceneKeys keys;  // content keys of the resident scene
    PackedScene pack;  // host staging of the last upload
    uint64_t hist_key = 0;
    std::vector<int32_t> prev_bvh_tri, prev_topo;
    int refits = 0;
    int last_refit = 0;
    float scene_ms[2] = {0.0f, 0.0f};  // last rm_set_scene: copies, refit (profiling on)

    // Frame buffers.
    GBufSet gbuf, hist_gbuf;
    ResSet rinit, rmerged, hist_res;
    DevBuf shade, den_a, hist_color;  // float4 per pixel
    DevBuf rgb8, hdr3, mega_hdr;
    DevBuf counters;
    // Pinned staging of each frame's read-back (counters, then the RGB8
    // image), copied to the caller's buffers after the frame's one wait.
    uint8_t* read_host = nullptr;
    size_t read_cap = 0;
    DevBuf queue;  // kQueueWords work counters of the persistent passes (band counters, shade order counts)
    DevBuf shade_order;  // restir_shade work order, one pixel index per entry
    // Wavefront schedule: two slot-state buffers, hits, control words, and a
    // pinned read-back of the active count with its events.
    DevBuf wave_state[2], wave_hits, wave_ctl;
    unsigned* wave_host = nullptr;
    hipEvent_t wave_ev[2] = {};
    size_t wave_cap = 0;
    DevBuf probe_o, probe_d, probe_t, probe_tri, probe_tuv;
    // Standalone merge / denoise passes on caller data (rm_restir_merge,
    // rm_denoise_image); separate from the frame buffers so the rendering
    // history is left as it was.
    GBufSet x_gbuf, x_hgbuf;
    ResSet x_rinit, x_rout, x_hres;
    DevBuf x_col, x_hcol, x_a, x_b;
    DevBuf adj_img, adj_tri_group, adj_light_group, adj_base, adj_grad, adj_bad, adj_primal, adj_record;
    std::vector<int32_t> light_tri;  // resident lights' triangles, CDF order
    // rm_set_group_materials: the material table with one slot per (group,
    // packed material) pair, each group's slots, every light's group and the
    // light records, built for the tri_group of key grp_key (0: none yet; any
    // scene upload resets it).
    uint64_t grp_key = 0;
    std::vector<std::vector<int32_t>> grp_slots;
    std::vector<F4> grp_mats, grp_lights;
    std::vector<int32_t> grp_light;

    // Learned denoiser: parameters (fp32 master, bf16 copy), gradient, Adam
    // moments, flipped weights for the input gradients, activations of the
    // last forward pass and its scratch, and the device-resident training set.
    bool kpn_ready = false;
    int kpn_step = 0;
    DevBuf kpn_p, kpn_pb, kpn_g, kpn_m, kpn_v, kpn_wd;
    DevBuf kpn_x[kKpnLayers], kpn_logits, kpn_dz[2], kpn_dlogit, kpn_partial, kpn_loss;
    DevBuf kpn_feat32, kpn_rad, kpn_ref, kpn_out, kpn_var, kpn_idx;
    DevBuf ds_feat, ds_rad, ds_ref;
    int ds_count = 0, ds_w = 0, ds_h = 0;

    // Temporal history (metal.mm:36-40).
    rm_camera hist_cam{};
    int hist_w = 0, hist_h = 0;
    uint64_t hist_seed = 0, hist_hash = 0;
    bool hist_valid = false;

    rm_stats stats{};

    // rm_trace_device: the device's scratch array and a pinned copy of its counters.
    unsigned long long* scratch_dev = nullptr;
    unsigned long long* scratch_host = nullptr;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    DevScene dev_scene() const {
        DevScene s{};
        s.nodes = nodes.as<float4>();
        s.geo = geo.as<float4>();
        s.nrm = nrm.as<float4>();
        s.nn = nn.as<float4>();
        s.matid = matid.as<int32_t>();
        s.mats = mats.as<float4>();
        s.uv = uv.as<float4>();
        s.lights = lights.as<float4>();
        s.cdf = cdf.as<float>();
        s.texels = texels.as<float>();
        s.nnodes = nnodes;
        s.nlights = nlights;
        s.opaque = opaque;
        s.light_total = light_total;
        return s;
    }
    void drop_history() { hist_valid = false; }
};

namespace {

int check(rm_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return RM_OK;
    c->drop_history();
    return c->fail(RM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Per-pass counters into rm_stats; index 5 holds the totals.
void collect(rm_stats& st, const unsigned long long* cnt) {
    st.rays[5] = st.node_visits[5] = st.tri_tests[5] = 0;
    for (int p = 0; p < kNumPasses; ++p) {
        st.rays[p] = cnt[3 * p];
        st.node_visits[p] = cnt[3 * p + 1];
        st.tri_tests[p] = cnt[3 * p + 2];
        st.rays[5] += st.rays[p];
        st.node_visits[5] += st.node_visits[p];
        st.tri_tests[5] += st.tri_tests[p];
    }
}

FrameParams frame_params(const rm_frame& f, uint64_t seed) {
    FrameParams P{};
    P.cam = f.camera;
    P.width = f.width; P.height = f.height; P.spp = f.spp; P.bounces = f.bounces;
    P.adapt_min = f.adapt_min;
    P.adapt_step = f.adapt_step > 0 ? f.adapt_step : 8;
    P.nomis = f.nomis;
    P.resolution = f.resolution; P.clamp_value = f.clamp_value;
    P.adapt_rel = f.adapt_rel; P.adapt_abs = f.adapt_abs;
    P.restir_candidates = f.restir_candidates; P.restir_spatial = f.restir_spatial;
    P.restir_mcap = f.restir_mcap; P.restir_radius = f.restir_radius;
    P.seed = seed;
    return P;
}

// Branch choice (metal.mm:232-254), with the reference's record sizes.
bool uses_restir(const rm_ctx* c, const FrameParams& P) {
    const size_t px = size_t(P.width) * size_t(P.height);
    const size_t restir_bytes = px * (sizeof(rm_gbuffer) + 2 * sizeof(rm_reservoir));
    return P.restir_spatial > 0 && c->nlights > 0 && P.bounces > 0 && c->opaque &&
           restir_bytes <= size_t(512) * 1024 * 1024;
}

// Megakernel lanes per pixel (metal.mm:390-394).
int megakernel_lanes(const FrameParams& P) {
    int lanes = 1;
    if (P.spp >= 8 && (P.adapt_min <= 0 || P.adapt_step % 8 == 0)) lanes = 8;
    if (P.spp >= 32 && (P.adapt_min <= 0 || P.adapt_step % 32 == 0)) lanes = 32;
    return lanes;
}

template <class V>
bool upload(rm_ctx* c, DevBuf& b, const V& v) {
    using T = typename V::value_type;
    if (!b.ensure(v.size() * sizeof(T))) return false;
    if (v.empty()) return true;
    return hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream) == hipSuccess;
}

// Packs the triangles' geometry (pack_geometry) and copies it to the device
// in kGeometrySlices ranges: each range goes to the device from the pinned
// packing buffers while the next one is packed. geo_dev receives the vertex
// records only when `positions`; nrm and nn always.
constexpr size_t kGeometrySlices = 8;
bool pack_upload_geometry(rm_ctx* c, const rm_scene_desc& s, PackedScene& ps, DevBuf& geo_dev, bool positions) {
    const size_t n = s.ntri;
    if (!geo_dev.ensure(n * kTriGeoWords * sizeof(F4)) || !c->nrm.ensure(n * kTriNrmWords * sizeof(F4)) ||
        !c->nn.ensure(n * sizeof(F4)))
        return false;
    auto copy = [&](DevBuf& d, const F4* src, size_t per_tri, size_t first, size_t count) {
        const size_t bytes = per_tri * sizeof(F4);
        return hipMemcpyAsync(static_cast<char*>(d.ptr) + first * bytes, src + first * per_tri, count * bytes,
                              hipMemcpyHostToDevice, c->stream) == hipSuccess;
    };
    return pack_geometry_sliced(s, ps, kGeometrySlices, [&](size_t first, size_t count) {
        return (!positions || copy(geo_dev, ps.geo.data(), kTriGeoWords, first, count)) &&
               copy(c->nrm, ps.nrm.data(), kTriNrmWords, first, count) && copy(c->nn, ps.nn.data(), 1, first, count);
    });
}

// Wavefront buffers for `pixels` slots.
bool wave_ensure(rm_ctx* c, size_t pixels, WaveBufs& wb) {
    const size_t bytes = wave_slot_words(true) * pixels * sizeof(uint32_t);
    if (!c->wave_state[0].ensure(bytes) || !c->wave_state[1].ensure(bytes) ||
        !c->wave_hits.ensure(pixels * sizeof(uint4)) || !c->wave_ctl.ensure(4 * sizeof(unsigned)))
        return false;
    if (!c->wave_host) {
        void* h = nullptr;
        if (hipHostMalloc(&h, 2 * sizeof(unsigned), hipHostMallocDefault) != hipSuccess) return false;
        c->wave_host = static_cast<unsigned*>(h);
    }
    for (auto& e : c->wave_ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            return false;
        }
    wb.state[0] = c->wave_state[0].as<uint32_t>();
    wb.state[1] = c->wave_state[1].as<uint32_t>();
    wb.hits = c->wave_hits.as<uint4>();
    wb.ctl = c->wave_ctl.as<unsigned>();
    wb.host = c->wave_host;
    wb.ev[0] = c->wave_ev[0];
    wb.ev[1] = c->wave_ev[1];
    wb.stride = pixels;
    wb.tail = c->schedule != 2;
    return true;
}

// Path engine pass (restir_shade or render_paths with one lane per pixel) on
// the context's schedule.
int path_pass(rm_ctx* c, bool shade, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb,
              const ReservoirSoA& rm, float4* hdr, uint8_t* rgb, float* lum_var, unsigned long long* cnt,
              rm_stats& st) {
    if (c->schedule == 0) {
        unsigned* order = nullptr;
        if (shade) {
            if (!c->shade_order.ensure(size_t(P.width) * size_t(P.height) * sizeof(unsigned))) {
                c->drop_history();
                return c->fail(RM_ERR_ALLOC, "HIP work-order buffer allocation failed");
            }
            order = c->shade_order.as<unsigned>();
        }
        launch_wave(c->stack_need, shade, sc, P, gb, rm, hdr, rgb, lum_var, c->queue.as<unsigned>() + kQueuePath,
                    order, c->queue.as<unsigned>() + kQueueOrder, cnt, c->stream);
        return RM_OK;
    }
    WaveBufs wb{};
    if (!wave_ensure(c, size_t(P.width) * size_t(P.height), wb)) {
        c->drop_history();
        return c->fail(RM_ERR_ALLOC, "HIP wavefront buffer allocation failed");
    }
    WaveRun wr{};
    const int e = run_wave(c->stack_need, shade, sc, P, gb, rm, hdr, rgb, lum_var, wb, cnt, c->stream, &wr);
    st.wave_iters = wr.iters;
    st.wave_tail = wr.tail;
    return check(c, hipError_t(e), shade ? "restir_shade" : "render_paths");
}

// Activation buffers for a forward (and backward) pass over `pixels` pixels.
bool kpn_ensure(rm_ctx* c, size_t pixels, bool train) {
    const KpnNet net = kpn_net();
    bool ok = c->kpn_x[0].ensure(pixels * kKpnFeat * 2) && c->kpn_logits.ensure(pixels * 32 * sizeof(float));
    for (int l = 1; l < kKpnLayers && ok; ++l) ok = c->kpn_x[l].ensure(pixels * kKpnHidden * 2);
    if (ok && train) {
        size_t scratch = 0;
        for (int l = 0; l < kKpnLayers; ++l)
            scratch = std::max(scratch, kpn_wgrad_scratch(net.layer[l].cin, net.layer[l].cout, int(pixels)));
        ok = c->kpn_dz[0].ensure(pixels * kKpnHidden * 2) && c->kpn_dz[1].ensure(pixels * kKpnHidden * 2) &&
             c->kpn_dlogit.ensure(pixels * 32 * 2) && c->kpn_partial.ensure(scratch * sizeof(float)) &&
             c->kpn_wd.ensure(size_t(kKpnHidden) * 9 * kKpnHidden * 2) && c->kpn_loss.ensure(sizeof(float));
    }
    return ok;
}

// Forward pass over n images of h x w: features kpn_x[0] (bf16), noisy
// radiance `rad`; with `ref`, the loss is added to *loss and the logits'
// gradient is left in kpn_dlogit.
void kpn_forward(rm_ctx* c, int n, int h, int w, const float* rad, const float* ref, float* out, float4* out4,
                 float* loss) {
    const KpnNet net = kpn_net();
    const __bf16* pb = c->kpn_pb.as<__bf16>();
    const float* p = c->kpn_p.as<float>();
    // An image alone (no loss, no [P][3] output): the head's logits go
    // straight into the filter inside the head convolution.
    const bool fused = !ref && !out && out4;
    for (int l = 0; l < kKpnLayers; ++l) {
        const KpnLayer& L = net.layer[l];
        const bool last = l == kKpnLayers - 1;
        if (last && fused) {
            launch_conv(L.cin, L.cout, 3, c->kpn_x[l].as<__bf16>(), pb + L.w_off, p + L.b_off, nullptr, out4, n, h, w,
                        c->stream, rad);
            return;
        }
        launch_conv(L.cin, L.cout, last ? 1 : 0, c->kpn_x[l].as<__bf16>(), pb + L.w_off, p + L.b_off, nullptr,
                    last ? c->kpn_logits.ptr : c->kpn_x[l + 1].ptr, n, h, w, c->stream);
    }
    const double count = double(n) * h * w * 3;
    launch_kpn_apply(c->kpn_logits.as<float>(), rad, ref, out, out4, ref ? c->kpn_dlogit.as<__bf16>() : nullptr, loss,
                     float(1.0 / count), n, h, w, c->stream);
}

// Gradient of the last training forward pass into kpn_g (every entry written).
void kpn_backward(rm_ctx* c, int n, int h, int w) {
    const KpnNet net = kpn_net();
    const __bf16* pb = c->kpn_pb.as<__bf16>();
    float* g = c->kpn_g.as<float>();
    (void)hipMemsetAsync(g, 0, net.nparam * sizeof(float), c->stream);  // padded weight entries stay 0
    const __bf16* dz = c->kpn_dlogit.as<__bf16>();
    int flip = 0;
    for (int l = kKpnLayers - 1; l >= 0; --l) {
        const KpnLayer& L = net.layer[l];
        launch_wgrad(L.cin, L.cout, c->kpn_x[l].as<__bf16>(), dz, c->kpn_partial.as<float>(), g + L.w_off,
                     g + L.b_off, n, h, w, c->stream);
        if (l == 0) break;
        // dL/d(pre-activation of layer l - 1) = (dz conv flipped weights) where x[l] > 0
        launch_flip_weights(L.cin, L.cout, pb + L.w_off, c->kpn_wd.as<__bf16>(), c->stream);
        __bf16* next = c->kpn_dz[flip].as<__bf16>();
        launch_conv(L.cout, L.cin, 2, dz, c->kpn_wd.as<__bf16>(), nullptr, c->kpn_x[l].as<__bf16>(), next, n, h, w,
                    c->stream);
        dz = next;
        flip ^= 1;
    }
}

}  // namespace

extern "C" {

int rm_create(rm_ctx** out, int hip_device) {
    std::lock_guard<std::mutex> lock(g_create_mu);
    if (!out) {
        g_create_error = "rm_create: null output pointer";
        return RM_ERR_INVALID;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        g_create_error = "No HIP GPU found";
        return RM_ERR_NO_DEVICE;
    }
    if (hip_device < 0 || hip_device >= count) {
        g_create_error = "HIP device index out of range";
        return RM_ERR_INVALID;
    }
    if (hipSetDevice(hip_device) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return RM_ERR_DEVICE;
    }
    rm_ctx* c = new rm_ctx();
    c->device = hip_device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        g_create_error = "Could not create HIP stream";
        return RM_ERR_DEVICE;
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            c->profiling = false;
        }
    }
    if (!c->counters.ensure(kNumCounters * sizeof(unsigned long long)) || !c->queue.ensure(kQueueWords * sizeof(unsigned))) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        g_create_error = "HIP counter allocation failed";
        return RM_ERR_ALLOC;
    }
    // The scratch array's address on this device (looked up once per device),
    // and pinned host memory for its read-back.
    if (g_scratch_dev.size() < size_t(count)) g_scratch_dev.resize(size_t(count), nullptr);
    if (!g_scratch_dev[size_t(hip_device)]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            rm_destroy(c);
            g_create_error = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
            return RM_ERR_DEVICE;
        }
        g_scratch_dev[size_t(hip_device)] = p;
    }
    c->scratch_dev = static_cast<unsigned long long*>(g_scratch_dev[size_t(hip_device)]);
    void* hp = nullptr;
    if (hipHostMalloc(&hp, kScratchWords * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        rm_destroy(c);
        g_create_error = "HIP pinned allocation failed";
        return RM_ERR_ALLOC;
    }
    c->scratch_host = static_cast<unsigned long long*>(hp);
    g_create_error.clear();
    *out = c;
    return RM_OK;
}

void rm_destroy(rm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->nodes, &c->geo, &c->nrm, &c->nn, &c->matid, &c->mats, &c->uv, &c->lights, &c->cdf,
                      &c->texels, &c->level_nodes, &c->shade, &c->den_a, &c->hist_color, &c->rgb8, &c->hdr3,
                      &c->mega_hdr, &c->counters, &c->queue, &c->shade_order, &c->probe_o, &c->probe_d,
                      &c->probe_t, &c->probe_tri, &c->probe_tuv, &c->adj_img, &c->adj_tri_group, &c->adj_light_group, &c->adj_base,
                      &c->adj_grad, &c->adj_bad};
    for (DevBuf* b : bufs) b->release();
    DevBuf* kbufs[] = {&c->kpn_p, &c->kpn_pb, &c->kpn_g, &c->kpn_m, &c->kpn_v, &c->kpn_wd, &c->kpn_logits,
                       &c->kpn_dz[0], &c->kpn_dz[1], &c->kpn_dlogit, &c->kpn_partial, &c->kpn_loss, &c->kpn_feat32,
                       &c->kpn_rad, &c->kpn_ref, &c->kpn_out, &c->kpn_var, &c->kpn_idx, &c->ds_feat, &c->ds_rad,
                       &c->ds_ref};
    for (DevBuf* b : kbufs) b->release();
    for (auto& b : c->kpn_x) b.release();
    c->gbuf.release();
    c->hist_gbuf.release();
    c->rinit.release();
    c->rmerged.release();
    c->hist_res.release();
    c->x_gbuf.release();
    c->x_hgbuf.release();
    c->x_rinit.release();
    c->x_rout.release();
    c->x_hres.release();
    c->x_col.release();
    c->x_hcol.release();
    c->x_a.release();
    c->x_b.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    c->geo_scene.release();
    c->geo_stage.release();
    c->nrm_stage.release();
    c->nn_stage.release();
    c->perm.release();
    c->build_scratch.release();
    if (c->build_host) (void)hipHostFree(c->build_host);
    for (auto& b : c->wave_state) b.release();
    c->wave_hits.release();
    c->wave_ctl.release();
    for (auto& e : c->wave_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->wave_host) (void)hipHostFree(c->wave_host);
    if (c->read_host) (void)hipHostFree(c->read_host);
    if (c->scratch_host) (void)hipHostFree(c->scratch_host);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* rm_last_error(const rm_ctx* c) {
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

int rm_set_profiling(rm_ctx* c, int enable) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    for (auto& e : c->ev)
        if (enable && !e) return c->fail(RM_ERR_DEVICE, "HIP events unavailable for profiling");
    c->profiling = enable != 0;
    return RM_OK;
}

int rm_set_schedule(rm_ctx* c, int schedule) {
    if (!c || schedule < 0 || schedule > 2) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    c->schedule = schedule;
    return RM_OK;
}

int rm_get_stats(rm_ctx* c, rm_stats* out) {
    if (!c || !out) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    *out = c->stats;
    return RM_OK;
}

// The records' keys: from the chunk values the caller took while writing
// them (pre), else hashed here.
static SceneKeys keys_of(const rm_scene_desc& s, const rmkeys::RecordKeys* pre) {
    if (pre && pre->chunks == rmkeys::key_chunks(s.ntri) && (pre->chunks == 0 || (pre->kp && pre->kn && pre->km)))
        return scene_keys_from_chunks(s, *pre);
    return scene_keys(s);
}

static int set_scene(rm_ctx* c, const rm_scene_desc* s, const rmkeys::RecordKeys* pre) {
    if (!c || !s) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    std::string perr;
    if (!validate_scene(*s, false, false, false, perr)) return c->fail(RM_ERR_INVALID, perr);
    const SceneKeys keys = keys_of(*s, pre);
    const bool had = c->has_scene;
    if (had && keys == c->keys && !c->gpu_tree) {
        c->last_refit = 0;
        c->stats.refit = 0;
        return RM_OK;  // identical scene already resident
    }
    const bool same_count = had && keys.ntri == c->keys.ntri;
    const bool topo_same = same_count && !c->gpu_tree && same_topology(*s, c->prev_bvh_tri, c->prev_topo);
    const bool moved = !same_count || keys.positions != c->keys.positions;
    const bool normals = !same_count || keys.normals != c->keys.normals;
    const bool materials = !same_count || keys.material != c->keys.material || keys.ntexel != c->keys.ntexel;
    // The tree follows the reference's acceleration-structure policy
    // (metal.mm:194-231): when positions changed and the topology is the same,
    // refit the resident 4-wide tree on the GPU, at most 31 times in a row;
    // otherwise rebuild it from the caller's nodes.
    const bool rebuild = !topo_same || (moved && c->refits >= 31);
    const bool refit = !rebuild && moved;
    const bool geometry = moved || normals || !topo_same;  // leaf order follows the tree
    if (!validate_scene(*s, !topo_same, materials, true, perr)) return c->fail(RM_ERR_INVALID, perr);

    PackedScene& ps = c->pack;  // reused: an animated frame re-packs into the same host buffers
    if (rebuild) {
        pack_tree(*s, ps);
        if (ps.stack_need > kMaxStack)
            return c->fail(RM_ERR_INVALID, "BVH too deep: the traversal stack would need " +
                                               std::to_string(ps.stack_need) + " entries (at most " +
                                               std::to_string(kMaxStack) + ")");
    }
    c->has_scene = false;  // until every part is resident again
    const bool prof = c->profiling;
    if (prof) (void)hipEventRecord(c->ev[8], c->stream);
    bool ok = !geometry || pack_upload_geometry(c, *s, ps, c->geo, true);
    pack_lights(*s, ps);
    if (materials) pack_materials(*s, ps);
    ok = ok && upload(c, c->lights, ps.lights) && upload(c, c->cdf, ps.cdf);
    if (ok && materials)
        ok = upload(c, c->matid, ps.matid) && upload(c, c->mats, ps.mats) && upload(c, c->uv, ps.uv) &&
             upload(c, c->texels, ps.texels);
    std::vector<int32_t> order;
    if (ok && rebuild) {
        ok = upload(c, c->nodes, ps.nodes);
        // Nodes grouped by depth, deepest level first, for the refit launches.
        const int nn4 = int(ps.depth.size());
        int maxd = 0;
        for (int d : ps.depth) maxd = std::max(maxd, d);
        std::vector<int32_t> count(nn4 > 0 ? size_t(maxd) + 2 : 1, 0);
        for (int d : ps.depth) ++count[size_t(maxd - d) + 1];
        for (size_t l = 1; l < count.size(); ++l) count[l] += count[l - 1];
        c->level_start.assign(count.begin(), count.end());
        order.resize(size_t(nn4));
        std::vector<int32_t> fill(count.begin(), count.end() - 1);
        for (int i = 0; i < nn4; ++i) order[size_t(fill[size_t(maxd - ps.depth[size_t(i)])]++)] = i;
        ok = ok && upload(c, c->level_nodes, order);
    }
    if (!ok) {
        c->drop_history();
        return c->fail(RM_ERR_ALLOC, "HIP scene buffer allocation failed");
    }
    if (prof) (void)hipEventRecord(c->ev[9], c->stream);
    if (refit)
        for (size_t l = 0; l + 1 < c->level_start.size(); ++l)
            launch_refit_level(c->nodes.as<float4>(), c->geo.as<float4>(),
                               c->level_nodes.as<int32_t>() + c->level_start[l],
                               c->level_start[l + 1] - c->level_start[l], c->stream);
    if (prof) (void)hipEventRecord(c->ev[10], c->stream);
    if (int rc = check(c, hipStreamSynchronize(c->stream), "scene upload")) return rc;
    c->scene_ms[0] = c->scene_ms[1] = 0.0f;
    if (prof) {
        if (hipEventElapsedTime(&c->scene_ms[0], c->ev[8], c->ev[9]) != hipSuccess) c->scene_ms[0] = 0.0f;
        if (hipEventElapsedTime(&c->scene_ms[1], c->ev[9], c->ev[10]) != hipSuccess) c->scene_ms[1] = 0.0f;
    }
    if (rebuild) {
        c->nnodes = int32_t(ps.depth.size());
        c->stack_need = ps.stack_need;
        c->refits = 0;
        c->gpu_tree = false;
        topology_of(*s, c->prev_bvh_tri, c->prev_topo);
    }
    if (refit) ++c->refits;
    if (materials) c->opaque = ps.opaque;
    c->nlights = int32_t(s->nlight);
    c->light_tri.assign(s->light_tri, s->light_tri + s->nlight);
    c->light_total = ps.light_total;
    c->ntri = int32_t(s->ntri);
    c->keys = keys;
    c->grp_key = 0;
    c->hist_key = history_key(keys, s->light_total);
    c->has_scene = true;
    c->last_refit = refit ? 1 : rebuild ? 2 : 0;
    c->stats.refit = c->last_refit;
    c->stats.bvh_nodes = c->nnodes;
    c->stats.stack_depth = c->stack_need;
    c->stats.scene_ms[0] = c->scene_ms[0];
    c->stats.scene_ms[1] = c->scene_ms[1];
    return RM_OK;
}

int rm_set_scene(rm_ctx* c, const rm_scene_desc* s) { return set_scene(c, s, nullptr); }

static int set_scene_gpu_tree(rm_ctx* c, const rm_scene_desc* s_in, int rebuild_request,
                              const rmkeys::RecordKeys* pre, bool staged) {
    if (!c || !s_in) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    rm_scene_desc sd = *s_in;  // the caller's tree, if any, is not used
    sd.nodes = nullptr;
    sd.nnode = 0;
    sd.bvh_tri = nullptr;
    const rm_scene_desc* s = &sd;
    std::string perr;
    if (!validate_scene(*s, false, false, false, perr)) return c->fail(RM_ERR_INVALID, perr);
    if (s->ntri == 0 || s->ntri > size_t(INT32_MAX / 4)) return c->fail(RM_ERR_INVALID, "GPU tree build needs 1 to 2^29 triangles");
    const SceneKeys keys = keys_of(*s, pre);
    const bool had = c->has_scene && c->gpu_tree;
    // staged geometry is taken by this call or dropped
    const bool use_stage = staged && c->stage_ok.load() && c->stage_ntri == s->ntri && c->staged_tris.load() == s->ntri;
    c->stage_ntri = 0;
    if (had && keys == c->keys && !rebuild_request) {
        c->last_refit = 0;
        c->stats.refit = 0;
        return RM_OK;
    }
    const bool same_count = had && keys.ntri == c->keys.ntri;
    const bool moved = !same_count || keys.positions != c->keys.positions;
    const bool normals = !same_count || keys.normals != c->keys.normals;
    const bool materials = !same_count || keys.material != c->keys.material || keys.ntexel != c->keys.ntexel;
    // Rebuild on request, when the triangle count changed, or after 31
    // refits in a row (the policy rm_set_scene applies to host trees).
    const bool rebuild = rebuild_request || !same_count || (moved && c->refits >= 31);
    const bool refit = !rebuild && moved;
    if (!validate_scene(*s, false, materials, true, perr)) return c->fail(RM_ERR_INVALID, perr);

    PackedScene& ps = c->pack;
    const int n = int(s->ntri);
    const size_t max_nodes = size_t(std::max(n - 1, 1));

    c->has_scene = false;
    const bool prof = c->profiling;
    if (prof) (void)hipEventRecord(c->ev[8], c->stream);
    // no tree: records in scene order (the staged copies already sit in the
    // staging buffers, queued on c->stream ahead of everything below)
    bool ok = true;
    if (moved || normals) {
        if (use_stage) {
            std::swap(c->geo_scene, c->geo_stage);
            std::swap(c->nrm, c->nrm_stage);
            std::swap(c->nn, c->nn_stage);
        } else {
            ok = pack_upload_geometry(c, *s, ps, c->geo_scene, moved);
        }
    }
    pack_lights(*s, ps);
    if (materials) pack_materials(*s, ps);
    ok = ok && upload(c, c->lights, ps.lights) && upload(c, c->cdf, ps.cdf);
    if (ok && materials)
        ok = upload(c, c->matid, ps.matid) && upload(c, c->mats, ps.mats) && upload(c, c->uv, ps.uv) &&
             upload(c, c->texels, ps.texels);
    if (ok && (rebuild || refit)) ok = c->geo.ensure(size_t(n) * kTriGeoWords * sizeof(float4));
    if (ok && rebuild)
        ok = c->perm.ensure(size_t(n) * sizeof(uint32_t)) && c->nodes.ensure(max_nodes * kNodeWords * sizeof(float4)) &&
             c->level_nodes.ensure(max_nodes * sizeof(int32_t)) && c->build_scratch.ensure(gpu_build_bytes(n));
    if (ok && rebuild && !c->build_host) {
        void* h = nullptr;
        ok = hipHostMalloc(&h, 2 * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess;
        c->build_host = static_cast<uint32_t*>(h);
    }
    if (!ok) {
        c->drop_history();
        return c->fail(RM_ERR_ALLOC, "HIP scene buffer allocation failed");
    }
    if (prof) (void)hipEventRecord(c->ev[9], c->stream);
    GpuBuildResult br;
    if (rebuild) {
        GpuBuildArgs a{};
        a.geo_scene = c->geo_scene.as<float4>();
        a.geo = c->geo.as<float4>();
        a.perm = c->perm.as<uint32_t>();
        a.nodes = c->nodes.as<float4>();
        a.level_nodes = c->level_nodes.as<int32_t>();
        a.max_nodes = max_nodes;
        a.scratch = c->build_scratch.ptr;
        a.scratch_bytes = c->build_scratch.cap;
        a.host_words = c->build_host;
        a.n = n;
        if (int rc = check(c, hipError_t(gpu_build(a, br, c->stream)), "GPU tree build")) return rc;
        if (br.stack_need > kMaxStack)
            return c->fail(RM_ERR_INVALID, "BVH too deep: the traversal stack would need " +
                                               std::to_string(br.stack_need) + " entries (at most " +
                                               std::to_string(kMaxStack) + ")");
        // Refit launch ranges into level_nodes (written by the build), deepest level first.
        c->level_start.assign(1, 0);
        for (size_t l = br.level_count.size(); l-- > 0;)
            c->level_start.push_back(c->level_start.back() + br.level_count[l]);
    } else if (refit) {
        launch_gather_tris(c->geo_scene.as<float4>(), c->perm.as<uint32_t>(), n, c->geo.as<float4>(), c->stream);
        for (size_t l = 0; l + 1 < c->level_start.size(); ++l)
            launch_refit_level(c->nodes.as<float4>(), c->geo.as<float4>(),
                               c->level_nodes.as<int32_t>() + c->level_start[l],
                               c->level_start[l + 1] - c->level_start[l], c->stream);
    }
    if (prof) (void)hipEventRecord(c->ev[10], c->stream);
    if (int rc = check(c, hipStreamSynchronize(c->stream), "scene upload")) return rc;
    c->scene_ms[0] = c->scene_ms[1] = 0.0f;
    if (prof) {
        if (hipEventElapsedTime(&c->scene_ms[0], c->ev[8], c->ev[9]) != hipSuccess) c->scene_ms[0] = 0.0f;
        if (hipEventElapsedTime(&c->scene_ms[1], c->ev[9], c->ev[10]) != hipSuccess) c->scene_ms[1] = 0.0f;
    }
    if (rebuild) {
        c->nnodes = br.nodes;
        c->stack_need = br.stack_need;
        c->refits = 0;
        c->gpu_tree = true;
        c->prev_bvh_tri.clear();
        c->prev_topo.clear();
    }
    if (refit) ++c->refits;
    if (materials) c->opaque = ps.opaque;
    c->nlights = int32_t(s->nlight);
    c->light_tri.assign(s->light_tri, s->light_tri + s->nlight);
    c->light_total = ps.light_total;
    c->ntri = n;
    c->keys = keys;
    c->grp_key = 0;
    c->hist_key = history_key(keys, s->light_total);
    c->has_scene = true;
    c->last_refit = refit ? 1 : rebuild ? 3 : 0;
    c->stats.refit = c->last_refit;
    c->stats.bvh_nodes = c->nnodes;
    c->stats.stack_depth = c->stack_need;
    c->stats.scene_ms[0] = c->scene_ms[0];
    c->stats.scene_ms[1] = c->scene_ms[1];
    return RM_OK;
}

int rm_set_scene_gpu_tree(rm_ctx* c, const rm_scene_desc* s, int rebuild_request) {
    return set_scene_gpu_tree(c, s, rebuild_request, nullptr, false);
}

}  // extern "C"

int rm_set_scene_keyed(rm_ctx* c, const rm_scene_desc* s, const rmkeys::RecordKeys& keys) {
    return set_scene(c, s, &keys);
}

int rm_set_scene_gpu_tree_keyed(rm_ctx* c, const rm_scene_desc* s, int rebuild_request,
                                const rmkeys::RecordKeys& keys, bool staged) {
    return set_scene_gpu_tree(c, s, rebuild_request, &keys, staged);
}

int rm_geo_stage_begin(rm_ctx* c, size_t ntri) {
    if (!c || ntri == 0 || ntri > size_t(INT32_MAX / 4)) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    c->stage_ntri = 0;
    if (!c->geo_stage.ensure(ntri * kTriGeoWords * sizeof(F4)) || !c->nrm_stage.ensure(ntri * kTriNrmWords * sizeof(F4)) ||
        !c->nn_stage.ensure(ntri * sizeof(F4)))
        return c->fail(RM_ERR_ALLOC, "HIP geometry staging allocation failed");
    c->stage_ntri = ntri;
    c->staged_tris.store(0);
    c->stage_ok.store(true);
    return RM_OK;
}

bool rm_geo_stage_put(rm_ctx* c, size_t first, size_t count, const F4* geo, const F4* nrm, const F4* nn) {
    if (!c || first > c->stage_ntri || count > c->stage_ntri - first || !geo || !nrm || !nn) {
        if (c) c->stage_ok.store(false);
        return false;
    }
    auto copy = [&](DevBuf& d, const F4* src, size_t per_tri) {
        const size_t bytes = per_tri * sizeof(F4);
        return hipMemcpyAsync(static_cast<char*>(d.ptr) + first * bytes, src, count * bytes, hipMemcpyHostToDevice,
                              c->stream) == hipSuccess;
    };
    // the calling thread may be a worker whose current device is another one
    const bool ok = hipSetDevice(c->device) == hipSuccess && copy(c->geo_stage, geo, kTriGeoWords) &&
                    copy(c->nrm_stage, nrm, kTriNrmWords) && copy(c->nn_stage, nn, 1);
    if (!ok) c->stage_ok.store(false);
    c->staged_tris.fetch_add(count);
    return ok;
}

namespace {

// Frame read-back through the context's pinned staging buffer: the counters
// at offset 0, the RGB8 image (rgb_bytes, may be 0) after them. A copy into
// pageable memory would be staged by the runtime with a wait of its own.
constexpr size_t kReadImageOffset = 256;
static_assert(kNumCounters * sizeof(unsigned long long) <= kReadImageOffset, "counters fit before the image");

int queue_readback(rm_ctx* c, size_t rgb_bytes, const unsigned long long* cnt) {
    const size_t need = kReadImageOffset + rgb_bytes;
    if (c->read_cap < need) {
        // no copy into the old buffer may still be in flight (a frame that
        // failed after queueing its read-back did not wait for it)
        if (int rc = check(c, hipStreamSynchronize(c->stream), "read-back")) return rc;
        if (c->read_host) (void)hipHostFree(c->read_host);
        c->read_host = nullptr;
        c->read_cap = 0;
        void* h = nullptr;
        if (hipHostMalloc(&h, need, hipHostMallocDefault) != hipSuccess) {
            c->drop_history();
            return c->fail(RM_ERR_ALLOC, "HIP pinned read-back allocation failed");
        }
        c->read_host = static_cast<uint8_t*>(h);
        c->read_cap = need;
    }
    if (int rc = check(c, hipMemcpyAsync(c->read_host, cnt, kNumCounters * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, c->stream), "counters"))
        return rc;
    if (rgb_bytes)
        if (int rc = check(c, hipMemcpyAsync(c->read_host + kReadImageOffset, c->rgb8.ptr, rgb_bytes,
                                             hipMemcpyDeviceToHost, c->stream), "readback"))
            return rc;
    return RM_OK;
}

// After the frame's wait: the counters to host_cnt, the image to rgb8 (if given).
void finish_readback(const rm_ctx* c, uint8_t* rgb8, size_t rgb_bytes, unsigned long long* host_cnt) {
    std::memcpy(host_cnt, c->read_host, kNumCounters * sizeof(unsigned long long));
    if (rgb8) std::memcpy(rgb8, c->read_host + kReadImageOffset, rgb_bytes);
}

// rm_render with the context locked; with feat_host / rad_host (ReSTIR frames
// only) also the denoiser's input features and noisy radiance.
int render_locked(rm_ctx* c, const rm_frame* f, uint64_t seed, uint8_t* rgb8, float* hdr, float* feat_host,
                  float* rad_host) {
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (f->width <= 0 || f->height <= 0 || f->spp <= 0 || f->bounces < 0 ||
        size_t(f->width) * size_t(f->height) > 100000000)
        return c->fail(RM_ERR_INVALID, "Invalid render dimensions or sample count");
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    const size_t px = size_t(f->width) * size_t(f->height);
    const int npx = int(px);

    FrameParams P = frame_params(*f, seed);
    const DevScene sc = c->dev_scene();
    const bool use_spatial = uses_restir(c, P);
    const bool learned = use_spatial && f->denoise == 2;
    const bool features = use_spatial && (learned || feat_host);
    if ((feat_host || rad_host) && !use_spatial)
        return c->fail(RM_ERR_INVALID, "denoiser features need a ReSTIR frame (restir_spatial > 0, lights, opaque)");
    if (learned && !c->kpn_ready) return c->fail(RM_ERR_INVALID, "denoise = 2 needs weights (rm_kpn_set_weights)");
    bool have_history = use_spatial && c->hist_valid && c->hist_w == f->width && c->hist_h == f->height &&
                        seed != c->hist_seed && c->hist_key == c->hist_hash;
    if (!f->temporal) have_history = false;
    P.has_history = have_history ? 1 : 0;
    P.prev = c->hist_cam;

    if (!c->rgb8.ensure(px * 3) || (hdr && !c->hdr3.ensure(px * 3 * sizeof(float)))) {
        c->drop_history();
        return c->fail(RM_ERR_ALLOC, "HIP image buffer allocation failed");
    }
    unsigned long long* cnt = c->counters.as<unsigned long long>();
    if (int rc = check(c, hipMemsetAsync(cnt, 0, kNumCounters * sizeof(unsigned long long), c->stream), "counters"))
        return rc;
    if (int rc = check(c, hipMemsetAsync(c->queue.ptr, 0, kQueueWords * sizeof(unsigned), c->stream), "work queue")) return rc;
    rm_stats st{};
    st.bvh_nodes = c->nnodes;
    st.stack_depth = c->stack_need;
    st.refit = c->last_refit;
    st.scene_ms[0] = c->scene_ms[0];
    st.scene_ms[1] = c->scene_ms[1];
    st.has_history = P.has_history;
    const bool prof = c->profiling;
    auto mark = [&](int i) {
        if (prof) (void)hipEventRecord(c->ev[i], c->stream);
    };
    mark(0);
    int denoise_passes = 0;
    if (use_spatial) {
        st.path = 1;
        if (!c->gbuf.ensure(px) || !c->hist_gbuf.ensure(px) || !c->rinit.ensure(px) || !c->rmerged.ensure(px) ||
            !c->hist_res.ensure(px) || !c->shade.ensure(px * sizeof(float4)) || !c->den_a.ensure(px * sizeof(float4)) ||
            !c->hist_color.ensure(px * sizeof(float4))) {
            c->drop_history();
            return c->fail(RM_ERR_ALLOC, "HIP ReSTIR buffer allocation failed");
        }
        const GBufferSoA gb = c->gbuf.soa(), hg = c->hist_gbuf.soa();
        const ReservoirSoA ri = c->rinit.soa(), rmg = c->rmerged.soa(), hr = c->hist_res.soa();
        launch_generate(c->stack_need, sc, P, gb, ri, c->queue.as<unsigned>() + kQueueGenerate, cnt, c->stream);
        mark(1);
        launch_spatial(c->stack_need, sc, P, gb, ri, rmg, hg, hr, cnt, c->stream);
        mark(2);
        if (features && (!c->kpn_var.ensure(px * sizeof(float)) || !c->kpn_feat32.ensure(px * kKpnFeat * 4) ||
                         !c->kpn_rad.ensure(px * 3 * sizeof(float)) || (learned && !kpn_ensure(c, px, false)))) {
            c->drop_history();
            return c->fail(RM_ERR_ALLOC, "HIP denoiser buffer allocation failed");
        }
        if (int rc = path_pass(c, true, sc, P, gb, rmg, c->shade.as<float4>(), nullptr,
                               features ? c->kpn_var.as<float>() : nullptr, cnt, st))
            return rc;
        mark(3);
        const float4* final_img = c->shade.as<float4>();
        bool final_in_hist = false;
        // Features: fp32 when the caller reads them back, bf16 straight into
        // the network's input for a learned frame.
        if (features)
            launch_features(P, gb, c->shade.as<float4>(), c->kpn_var.as<float>(),
                            feat_host ? c->kpn_feat32.as<float>() : nullptr, c->kpn_rad.as<float>(), c->stream,
                            learned ? c->kpn_x[0].as<__bf16>() : nullptr);
        if (learned) {
            // Kernel-predicting network in place of the a-trous filter.
            mark(4);
            kpn_forward(c, 1, f->height, f->width, c->kpn_rad.as<float>(), nullptr, nullptr, c->den_a.as<float4>(),
                        nullptr);
            mark(5);
            denoise_passes = 1;
            final_img = c->den_a.as<float4>();
        } else if (f->denoise) {
            // Temporal pass into A, then a-trous passes ping-ponging between A
            // and the history color buffer (its old content has been consumed).
            float4* a = c->den_a.as<float4>();
            float4* b = c->hist_color.as<float4>();
            launch_temporal(P, gb, hg, c->shade.as<float4>(), c->hist_color.as<float4>(), a, c->stream);
            mark(4);
            const int iters = std::max(0, f->denoise_iterations);
            for (int s = 0; s < iters; ++s) {
                launch_atrous(P, gb, a, b, s, c->stream);
                std::swap(a, b);
            }
            denoise_passes = iters;
            mark(5);
            final_img = a;
            final_in_hist = a == c->hist_color.as<float4>();
        } else {
            mark(4);
            mark(5);
        }
        launch_tonemap(npx, final_img, c->rgb8.as<uint8_t>(), hdr ? c->hdr3.as<float>() : nullptr, c->stream);
        mark(6);
        if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
        if (int rc = queue_readback(c, rgb8 ? px * 3 : 0, cnt)) return rc;
        if (hdr)
            if (int rc = check(c, hipMemcpyAsync(hdr, c->hdr3.ptr, px * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                                 c->stream), "readback"))
                return rc;
        if (feat_host)
            if (int rc = check(c, hipMemcpyAsync(feat_host, c->kpn_feat32.ptr, px * kKpnFeat * 4, hipMemcpyDeviceToHost,
                                                 c->stream), "readback"))
                return rc;
        if (rad_host)
            if (int rc = check(c, hipMemcpyAsync(rad_host, c->kpn_rad.ptr, px * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                                 c->stream), "readback"))
                return rc;
        if (int rc = check(c, hipStreamSynchronize(c->stream), "frame")) return rc;
        unsigned long long host_cnt[kNumCounters];
        finish_readback(c, rgb8, px * 3, host_cnt);
        if (host_cnt[kCntOverflow]) {
            c->drop_history();
            return c->fail(RM_ERR_DEVICE, "traversal stack overflow");
        }
        // Publish this frame as the next frame's history (metal.mm:374-385).
        std::swap(c->gbuf, c->hist_gbuf);
        std::swap(c->rmerged, c->hist_res);
        // The history colour is the frame's denoised output (metal.mm:378,
        // which swaps it after every denoised frame); a learned frame's output
        // lies in den_a as the a-trous chain's may.
        if (f->denoise != 0 && !final_in_hist) std::swap(c->den_a, c->hist_color);
        c->hist_cam = f->camera;
        c->hist_w = f->width;
        c->hist_h = f->height;
        c->hist_seed = seed;
        c->hist_hash = c->hist_key;
        c->hist_valid = true;
        collect(st, host_cnt);
    } else {
        // Megakernel with 1, 8 or 32 lanes per pixel (metal.mm:390-394).
        const int lanes = megakernel_lanes(P);
        st.path = 0;
        st.lanes = lanes;
        if (hdr && !c->mega_hdr.ensure(px * sizeof(float4))) {
            c->drop_history();
            return c->fail(RM_ERR_ALLOC, "HIP image buffer allocation failed");
        }
        if (lanes == 1) {
            if (int rc = path_pass(c, false, sc, P, GBufferSoA{}, ReservoirSoA{},
                                   hdr ? c->mega_hdr.as<float4>() : nullptr, c->rgb8.as<uint8_t>(), nullptr, cnt, st))
                return rc;
        } else
            launch_render_paths(c->stack_need, lanes, sc, P, c->rgb8.as<uint8_t>(),
                                hdr ? c->mega_hdr.as<float4>() : nullptr, cnt, c->stream);
        mark(1);
        if (hdr) launch_tonemap(npx, c->mega_hdr.as<float4>(), c->rgb8.as<uint8_t>(), c->hdr3.as<float>(), c->stream);
        if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
        if (int rc = queue_readback(c, rgb8 ? px * 3 : 0, cnt)) return rc;
        if (hdr)
            if (int rc = check(c, hipMemcpyAsync(hdr, c->hdr3.ptr, px * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                                 c->stream), "readback"))
                return rc;
        if (int rc = check(c, hipStreamSynchronize(c->stream), "frame")) return rc;
        unsigned long long host_cnt[kNumCounters];
        finish_readback(c, rgb8, px * 3, host_cnt);
        if (host_cnt[kCntOverflow]) {
            c->drop_history();
            return c->fail(RM_ERR_DEVICE, "traversal stack overflow");
        }
        collect(st, host_cnt);
    }
    if (prof) {
        auto ms = [&](int a, int b) {
            float v = 0.0f;
            if (hipEventElapsedTime(&v, c->ev[a], c->ev[b]) != hipSuccess) v = 0.0f;
            return v;
        };
        if (st.path == 1) {
            st.ms[1] = ms(0, 1);
            st.ms[2] = ms(1, 2);
            st.ms[3] = ms(2, 3);
            st.ms[4] = ms(3, 4);
            st.ms[5] = denoise_passes > 0 ? ms(4, 5) : 0.0f;
            st.ms[6] = ms(5, 6);
            st.ms[7] = ms(0, 6);
        } else {
            st.ms[0] = ms(0, 1);
            st.ms[7] = ms(0, 1);
        }
    }
    c->stats = st;
    c->err.clear();
    return RM_OK;
}

// Reference-layout records (76-B G-buffer, 60-B reservoir) to and from the
// device's SoA float4 arrays (GBufferSoA, ReservoirSoA).
void gbuf_to_soa(const rm_gbuffer* g, size_t px, std::vector<float4> (&o)[5]) {
    for (auto& v : o) v.resize(px);
    for (size_t i = 0; i < px; ++i) {
        float kind;
        std::memcpy(&kind, &g[i].kind, 4);
        o[0][i] = make_float4(g[i].pos.x, g[i].pos.y, g[i].pos.z, kind);
        o[1][i] = make_float4(g[i].normal.x, g[i].normal.y, g[i].normal.z, g[i].rough);
        o[2][i] = make_float4(g[i].albedo.x, g[i].albedo.y, g[i].albedo.z, g[i].metallic);
        o[3][i] = make_float4(g[i].wo.x, g[i].wo.y, g[i].wo.z, g[i].ior);
        o[4][i] = make_float4(g[i].emission.x, g[i].emission.y, g[i].emission.z, 0.0f);
    }
}

void res_to_soa(const rm_reservoir* r, size_t px, std::vector<float4> (&o)[4]) {
    for (auto& v : o) v.resize(px);
    for (size_t i = 0; i < px; ++i) {
        float m, valid;
        std::memcpy(&m, &r[i].m, 4);
        std::memcpy(&valid, &r[i].valid, 4);
        o[0][i] = make_float4(r[i].light_pos.x, r[i].light_pos.y, r[i].light_pos.z, r[i].p_hat);
        o[1][i] = make_float4(r[i].light_normal.x, r[i].light_normal.y, r[i].light_normal.z, r[i].alpha);
        o[2][i] = make_float4(r[i].emission.x, r[i].emission.y, r[i].emission.z, r[i].sum_ws);
        o[3][i] = make_float4(r[i].sum_wu, m, valid, 0.0f);
    }
}

void res_from_soa(const std::vector<float4> (&a)[4], size_t px, rm_reservoir* out) {
    for (size_t i = 0; i < px; ++i) {
        rm_reservoir& o = out[i];
        o.light_pos = {a[0][i].x, a[0][i].y, a[0][i].z};
        o.p_hat = a[0][i].w;
        o.light_normal = {a[1][i].x, a[1][i].y, a[1][i].z};
        o.alpha = a[1][i].w;
        o.emission = {a[2][i].x, a[2][i].y, a[2][i].z};
        o.sum_ws = a[2][i].w;
        o.sum_wu = a[3][i].x;
        std::memcpy(&o.m, &a[3][i].y, 4);
        std::memcpy(&o.valid, &a[3][i].z, 4);
    }
}

// Queues host-to-device copies of SoA arrays into a set's buffers.
template <size_t N>
bool upload_soa(rm_ctx* c, DevBuf* bufs, const std::vector<float4> (&v)[N]) {
    for (size_t k = 0; k < N; ++k)
        if (!upload(c, bufs[k], v[k])) return false;
    return true;
}

// Linear RGB triples to float4 per pixel.
std::vector<float4> rgb_to_float4(const float* rgb, size_t px) {
    std::vector<float4> v(px);
    for (size_t i = 0; i < px; ++i) v[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
    return v;
}

// Grow-only buffer that keeps its contents (the training set).
bool grow_keep(DevBuf& b, size_t bytes, hipStream_t st) {
    if (bytes <= b.cap) return true;
    DevBuf n;
    if (!n.ensure(std::max(bytes, 2 * b.cap))) return false;
    if (b.ptr && b.cap && hipMemcpyAsync(n.ptr, b.ptr, b.cap, hipMemcpyDeviceToDevice, st) != hipSuccess) return false;
    if (hipStreamSynchronize(st) != hipSuccess) return false;
    b.release();
    b = n;
    n.ptr = nullptr;
    n.cap = 0;
    return true;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Uploads n feature images (fp32, kKpnFeat per pixel) as bf16 into kpn_x[0],
// and their radiance (and reference) into kpn_rad (kpn_ref).
int kpn_upload(rm_ctx* c, const float* feat, const float* noisy, const float* ref, size_t px) {
    hipStream_t st = c->stream;
    if (!c->kpn_feat32.ensure(px * kKpnFeat * 4) || !c->kpn_rad.ensure(px * 12) || (ref && !c->kpn_ref.ensure(px * 12)))
        return c->fail(RM_ERR_ALLOC, "HIP denoiser buffer allocation failed");
    if (hipMemcpyAsync(c->kpn_feat32.ptr, feat, px * kKpnFeat * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(c->kpn_rad.ptr, noisy, px * 12, hipMemcpyHostToDevice, st) != hipSuccess ||
        (ref && hipMemcpyAsync(c->kpn_ref.ptr, ref, px * 12, hipMemcpyHostToDevice, st) != hipSuccess))
        return c->fail(RM_ERR_DEVICE, "denoiser upload failed");
    launch_to_bf16(c->kpn_feat32.as<float>(), c->kpn_x[0].as<__bf16>(), px * kKpnFeat, st);
    return RM_OK;
}

}  // namespace

extern "C" {

int rm_render(rm_ctx* c, const rm_frame* f, uint64_t seed, uint8_t* rgb8, float* hdr) {
    if (!c || !f) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    return render_locked(c, f, seed, rgb8, hdr, nullptr, nullptr);
}

int rm_render_features(rm_ctx* c, const rm_frame* f, uint64_t seed, float* features, float* noisy) {
    if (!c || !f || !features || !noisy) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    return render_locked(c, f, seed, nullptr, nullptr, features, noisy);
}

size_t rm_kpn_param_count(void) { return kpn_net().nparam; }

int rm_kpn_set_weights(rm_ctx* c, const float* params, size_t n) {
    if (!c || !params) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    const size_t np = kpn_net().nparam;
    if (n != np) return c->fail(RM_ERR_INVALID, "rm_kpn_set_weights: expected " + std::to_string(np) + " floats");
    const size_t bytes = np * sizeof(float);
    if (!c->kpn_p.ensure(bytes) || !c->kpn_pb.ensure(np * 2) || !c->kpn_g.ensure(bytes) || !c->kpn_m.ensure(bytes) ||
        !c->kpn_v.ensure(bytes))
        return c->fail(RM_ERR_ALLOC, "HIP denoiser parameter allocation failed");
    hipStream_t st = c->stream;
    if (hipMemcpyAsync(c->kpn_p.ptr, params, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(c->kpn_m.ptr, 0, bytes, st) != hipSuccess || hipMemsetAsync(c->kpn_v.ptr, 0, bytes, st) != hipSuccess)
        return c->fail(RM_ERR_DEVICE, "denoiser upload failed");
    launch_to_bf16(c->kpn_p.as<float>(), c->kpn_pb.as<__bf16>(), np, st);
    if (int rc = check(c, hipStreamSynchronize(st), "denoiser weights")) return rc;
    c->kpn_step = 0;
    c->kpn_ready = true;
    return RM_OK;
}

int rm_kpn_get_weights(rm_ctx* c, float* params, size_t n) {
    if (!c || !params) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (!c->kpn_ready) return c->fail(RM_ERR_INVALID, "no denoiser weights set");
    if (n != kpn_net().nparam) return c->fail(RM_ERR_INVALID, "rm_kpn_get_weights: wrong size");
    // hipMemcpy runs on the null stream, which does not wait for the context's stream.
    if (int rc = check(c, hipStreamSynchronize(c->stream), "readback")) return rc;
    if (int rc = check(c, hipMemcpy(params, c->kpn_p.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "readback"))
        return rc;
    return RM_OK;
}

int rm_kpn_dataset_clear(rm_ctx* c) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    c->ds_count = 0;
    return RM_OK;
}

int rm_kpn_dataset_add(rm_ctx* c, const float* features, const float* noisy, const float* reference, int width,
                       int height) {
    if (!c || !features || !noisy || !reference || width <= 0 || height <= 0) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (c->ds_count > 0 && (width != c->ds_w || height != c->ds_h))
        return c->fail(RM_ERR_INVALID, "training images must share one size");
    const size_t px = size_t(width) * height, at = size_t(c->ds_count) * px;
    hipStream_t st = c->stream;
    if (!grow_keep(c->ds_feat, (at + px) * kKpnFeat * 2, st) || !grow_keep(c->ds_rad, (at + px) * 12, st) ||
        !grow_keep(c->ds_ref, (at + px) * 12, st) || !c->kpn_feat32.ensure(px * kKpnFeat * 4))
        return c->fail(RM_ERR_ALLOC, "HIP training set allocation failed");
    if (hipMemcpyAsync(c->kpn_feat32.ptr, features, px * kKpnFeat * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(c->ds_rad.as<float>() + at * 3, noisy, px * 12, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(c->ds_ref.as<float>() + at * 3, reference, px * 12, hipMemcpyHostToDevice, st) != hipSuccess)
        return c->fail(RM_ERR_DEVICE, "training set upload failed");
    launch_to_bf16(c->kpn_feat32.as<float>(), c->ds_feat.as<__bf16>() + at * kKpnFeat, px * kKpnFeat, st);
    if (int rc = check(c, hipStreamSynchronize(st), "training set upload")) return rc;
    c->ds_w = width;
    c->ds_h = height;
    ++c->ds_count;
    return RM_OK;
}

int rm_kpn_train(rm_ctx* c, int steps, int batch, float lr, uint64_t seed, float* losses) {
    if (!c || steps < 0 || batch <= 0) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->kpn_ready) return c->fail(RM_ERR_INVALID, "no denoiser weights set");
    if (c->ds_count == 0) return c->fail(RM_ERR_INVALID, "empty training set");
    if (steps == 0) return RM_OK;
    const int h = c->ds_h, w = c->ds_w;
    const size_t ppi = size_t(h) * w, px = ppi * size_t(batch);
    hipStream_t st = c->stream;
    std::vector<int32_t> idx(size_t(steps) * batch);
    uint64_t s = seed;
    for (auto& v : idx) v = int32_t(splitmix(s) % uint64_t(c->ds_count));
    if (!kpn_ensure(c, px, true) || !c->kpn_rad.ensure(px * 12) || !c->kpn_ref.ensure(px * 12) ||
        !c->kpn_idx.ensure(idx.size() * 4) || !c->kpn_loss.ensure(size_t(steps) * sizeof(float)))
        return c->fail(RM_ERR_ALLOC, "HIP denoiser training allocation failed");
    if (hipMemcpyAsync(c->kpn_idx.ptr, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(c->kpn_loss.ptr, 0, size_t(steps) * sizeof(float), st) != hipSuccess)
        return c->fail(RM_ERR_DEVICE, "denoiser training upload failed");
    const size_t np = kpn_net().nparam;
    if (c->profiling) (void)hipEventRecord(c->ev[0], st);
    for (int k = 0; k < steps; ++k) {
        launch_gather(c->ds_feat.as<__bf16>(), c->ds_rad.as<float>(), c->ds_ref.as<float>(),
                      c->kpn_idx.as<int32_t>() + size_t(k) * batch, batch, int(ppi), c->kpn_x[0].as<__bf16>(),
                      c->kpn_rad.as<float>(), c->kpn_ref.as<float>(), st);
        kpn_forward(c, batch, h, w, c->kpn_rad.as<float>(), c->kpn_ref.as<float>(), nullptr, nullptr,
                    c->kpn_loss.as<float>() + k);
        kpn_backward(c, batch, h, w);
        launch_adam(c->kpn_p.as<float>(), c->kpn_g.as<float>(), c->kpn_m.as<float>(), c->kpn_v.as<float>(),
                    c->kpn_pb.as<__bf16>(), np, lr, 0.9f, 0.999f, 1e-8f, ++c->kpn_step, st);
    }
    if (c->profiling) (void)hipEventRecord(c->ev[1], st);
    if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
    if (losses)
        if (int rc = check(c, hipMemcpyAsync(losses, c->kpn_loss.ptr, size_t(steps) * sizeof(float),
                                             hipMemcpyDeviceToHost, st), "readback"))
            return rc;
    if (int rc = check(c, hipStreamSynchronize(st), "denoiser training")) return rc;
    rm_stats sts{};
    if (c->profiling && hipEventElapsedTime(&sts.ms[7], c->ev[0], c->ev[1]) != hipSuccess) sts.ms[7] = 0.0f;
    c->stats = sts;
    return RM_OK;
}

int rm_kpn_denoise(rm_ctx* c, const float* features, const float* noisy, int width, int height, float* out) {
    if (!c || !features || !noisy || !out || width <= 0 || height <= 0) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->kpn_ready) return c->fail(RM_ERR_INVALID, "no denoiser weights set");
    const size_t px = size_t(width) * height;
    if (!kpn_ensure(c, px, false) || !c->kpn_out.ensure(px * 12))
        return c->fail(RM_ERR_ALLOC, "HIP denoiser buffer allocation failed");
    if (int rc = kpn_upload(c, features, noisy, nullptr, px)) return rc;
    if (c->profiling) (void)hipEventRecord(c->ev[0], c->stream);
    kpn_forward(c, 1, height, width, c->kpn_rad.as<float>(), nullptr, c->kpn_out.as<float>(), nullptr, nullptr);
    if (c->profiling) (void)hipEventRecord(c->ev[1], c->stream);
    if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
    if (int rc = check(c, hipMemcpyAsync(out, c->kpn_out.ptr, px * 12, hipMemcpyDeviceToHost, c->stream), "readback"))
        return rc;
    if (int rc = check(c, hipStreamSynchronize(c->stream), "denoise")) return rc;
    rm_stats sts{};
    if (c->profiling && hipEventElapsedTime(&sts.ms[5], c->ev[0], c->ev[1]) != hipSuccess) sts.ms[5] = 0.0f;
    c->stats = sts;
    return RM_OK;
}

int rm_kpn_gradient(rm_ctx* c, const float* features, const float* noisy, const float* reference, int count, int width,
                    int height, float* grad, float* loss) {
    if (!c || !features || !noisy || !reference || !grad || count <= 0 || width <= 0 || height <= 0)
        return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->kpn_ready) return c->fail(RM_ERR_INVALID, "no denoiser weights set");
    const size_t px = size_t(width) * height * size_t(count);
    if (!kpn_ensure(c, px, true)) return c->fail(RM_ERR_ALLOC, "HIP denoiser buffer allocation failed");
    if (int rc = kpn_upload(c, features, noisy, reference, px)) return rc;
    if (int rc = check(c, hipMemsetAsync(c->kpn_loss.ptr, 0, sizeof(float), c->stream), "loss")) return rc;
    kpn_forward(c, count, height, width, c->kpn_rad.as<float>(), c->kpn_ref.as<float>(), nullptr, nullptr,
                c->kpn_loss.as<float>());
    kpn_backward(c, count, height, width);
    if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
    const size_t np = kpn_net().nparam;
    if (int rc = check(c, hipMemcpyAsync(grad, c->kpn_g.ptr, np * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                       "readback"))
        return rc;
    if (loss)
        if (int rc = check(c, hipMemcpyAsync(loss, c->kpn_loss.ptr, sizeof(float), hipMemcpyDeviceToHost, c->stream),
                           "readback"))
            return rc;
    return check(c, hipStreamSynchronize(c->stream), "denoiser gradient");
}

int rm_trace(rm_ctx* c, const float* origins, const float* dirs, const float* tmax, size_t n, int any, int32_t* tri,
             float* t, float* u, float* v) {
    if (!c || (n > 0 && (!origins || !dirs || !tri))) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    if (n == 0) return RM_OK;
    if (n > (size_t(1) << 30)) return c->fail(RM_ERR_INVALID, "too many rays in one call");
    if (!c->probe_o.ensure(n * 12) || !c->probe_d.ensure(n * 12) || (tmax && !c->probe_t.ensure(n * 4)) ||
        !c->probe_tri.ensure(n * 4) || !c->probe_tuv.ensure(n * 16))
        return c->fail(RM_ERR_ALLOC, "HIP ray buffer allocation failed");
    hipStream_t s = c->stream;
    unsigned long long* cnt = c->counters.as<unsigned long long>();
    if (int rc = check(c, hipMemsetAsync(cnt, 0, kNumCounters * sizeof(unsigned long long), s), "counters")) return rc;
    if (int rc = check(c, hipMemcpyAsync(c->probe_o.ptr, origins, n * 12, hipMemcpyHostToDevice, s), "ray upload"))
        return rc;
    if (int rc = check(c, hipMemcpyAsync(c->probe_d.ptr, dirs, n * 12, hipMemcpyHostToDevice, s), "ray upload"))
        return rc;
    if (tmax)
        if (int rc = check(c, hipMemcpyAsync(c->probe_t.ptr, tmax, n * 4, hipMemcpyHostToDevice, s), "ray upload"))
            return rc;
    FrameParams P{};
    const DevScene sc = c->dev_scene();
    if (c->profiling) (void)hipEventRecord(c->ev[0], s);
    launch_trace(c->stack_need, sc, P, int(n), c->probe_o.as<float>(), c->probe_d.as<float>(),
                 tmax ? c->probe_t.as<float>() : nullptr, any, c->probe_tri.as<int32_t>(), c->probe_tuv.as<float4>(),
                 cnt, s);
    if (c->profiling) (void)hipEventRecord(c->ev[1], s);
    if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
    std::vector<float> tuv(n * 4);
    unsigned long long host_cnt[kNumCounters];
    if (int rc = check(c, hipMemcpyAsync(tri, c->probe_tri.ptr, n * 4, hipMemcpyDeviceToHost, s), "readback")) return rc;
    if (int rc = check(c, hipMemcpyAsync(tuv.data(), c->probe_tuv.ptr, n * 16, hipMemcpyDeviceToHost, s), "readback"))
        return rc;
    if (int rc = check(c, hipMemcpyAsync(host_cnt, cnt, sizeof(host_cnt), hipMemcpyDeviceToHost, s), "counters"))
        return rc;
    if (int rc = check(c, hipStreamSynchronize(s), "trace")) return rc;
    if (host_cnt[kCntOverflow]) return c->fail(RM_ERR_DEVICE, "traversal stack overflow");
    for (size_t i = 0; i < n; ++i) {
        if (t) t[i] = tuv[4 * i];
        if (u) u[i] = tuv[4 * i + 1];
        if (v) v[i] = tuv[4 * i + 2];
    }
    rm_stats st{};
    collect(st, host_cnt);
    st.bvh_nodes = c->nnodes;
    st.stack_depth = c->stack_need;
    if (c->profiling && hipEventElapsedTime(&st.ms[7], c->ev[0], c->ev[1]) != hipSuccess) st.ms[7] = 0.0f;
    c->stats = st;
    return RM_OK;
}

int rm_set_group_materials(rm_ctx* c, const int32_t* tri_group, size_t ntri, int ngroup, const float* params,
                           const float* group_emission, const float* light_cdf, size_t nlight, float light_total) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    const PackedScene& ps = c->pack;
    if (!tri_group || !params || !group_emission || ngroup < 1 || ngroup > (1 << 20) || ntri != size_t(c->ntri) ||
        nlight != size_t(c->nlights) || (nlight && !light_cdf) || ps.matid.size() != ntri ||
        ps.lights.size() != nlight * kLightWords)
        return c->fail(RM_ERR_INVALID, "rm_set_group_materials: arguments do not match the resident scene");
    uint64_t key = 0xCBF29CE484222325ull ^ (uint64_t(ngroup) << 32) ^ ntri;
    for (size_t i = 0; i < ntri; ++i) {
        if (tri_group[i] < -1 || tri_group[i] >= ngroup)
            return c->fail(RM_ERR_INVALID, "rm_set_group_materials: group index out of range");
        key = (key ^ uint32_t(tri_group[i])) * 0x100000001B3ull;
    }
    if (key == 0) key = 1;
    if (key != c->grp_key) {
        // One slot per (group, packed material): a group's slots are written
        // in place by every later call with the same grouping.
        c->grp_mats.assign(ps.mats.begin(), ps.mats.end());
        c->grp_slots.assign(size_t(ngroup), {});
        std::vector<int32_t> matid(ntri);
        std::vector<std::vector<std::pair<int32_t, int32_t>>> seen(static_cast<size_t>(ngroup));  // (packed id, slot)
        for (size_t i = 0; i < ntri; ++i) {
            const int g = tri_group[i], orig = ps.matid[i];
            if (g < 0) {
                matid[i] = orig;
                continue;
            }
            int32_t slot = -1;
            for (const auto& pr : seen[size_t(g)])
                if (pr.first == orig) slot = pr.second;
            if (slot < 0) {
                slot = int32_t(c->grp_mats.size() / kMatWords);
                for (int w = 0; w < kMatWords; ++w) c->grp_mats.push_back(ps.mats[size_t(orig) * kMatWords + w]);
                seen[size_t(g)].push_back({orig, slot});
                c->grp_slots[size_t(g)].push_back(slot);
            }
            matid[i] = slot;
        }
        c->grp_light.resize(nlight);
        for (size_t k = 0; k < nlight; ++k) c->grp_light[k] = tri_group[c->light_tri[k]];
        c->grp_lights.assign(ps.lights.begin(), ps.lights.end());
        if (!upload(c, c->matid, matid)) return c->fail(RM_ERR_ALLOC, "HIP material buffer allocation failed");
        if (int rc = check(c, hipStreamSynchronize(c->stream), "material upload")) return rc;
        c->grp_key = key;
    }
    // Material rows [ar ag ab alpha] [er eg eb ior] [rough metallic shader texoff] [...] and
    // light records [...] [...] [...] [er eg eb alpha] of the groups.
    for (int g = 0; g < ngroup; ++g) {
        const float* p = params + 5 * size_t(g);
        const float er = group_emission[3 * g] * p[4], eg = group_emission[3 * g + 1] * p[4],
                    eb = group_emission[3 * g + 2] * p[4];
        for (int32_t slot : c->grp_slots[size_t(g)]) {
            F4* m = &c->grp_mats[size_t(slot) * kMatWords];
            m[0].x = p[0]; m[0].y = p[1]; m[0].z = p[2];
            m[1].x = er; m[1].y = eg; m[1].z = eb;
            m[2].x = p[3];
        }
    }
    for (size_t k = 0; k < nlight; ++k) {
        const int g = c->grp_light[k];
        if (g < 0) continue;
        const float* p = params + 5 * size_t(g);
        F4& e = c->grp_lights[4 * k + 3];
        e.x = group_emission[3 * g] * p[4];
        e.y = group_emission[3 * g + 1] * p[4];
        e.z = group_emission[3 * g + 2] * p[4];
    }
    const std::vector<float> cdf(light_cdf, light_cdf + nlight);
    if (!upload(c, c->mats, c->grp_mats) || !upload(c, c->lights, c->grp_lights) || !upload(c, c->cdf, cdf))
        return c->fail(RM_ERR_ALLOC, "HIP material buffer allocation failed");
    if (int rc = check(c, hipStreamSynchronize(c->stream), "material upload")) return rc;
    c->light_total = light_total;
    // The resident content no longer matches the last scene's keys: a later
    // rm_set_scene re-sends materials and lights, and no history carries over.
    c->keys.material = c->keys.lights = ~0ull;
    c->drop_history();
    c->err.clear();
    return RM_OK;
}

int rm_read_tree(rm_ctx* c, float* nodes, size_t node_cap, float* tris, size_t tri_cap, int32_t* counts) {
    if (!c || !counts) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    counts[0] = c->nnodes;
    counts[1] = c->ntri;
    counts[2] = c->stack_need;
    const size_t nb = size_t(c->nnodes) * kNodeWords * sizeof(float4), tb = size_t(c->ntri) * kTriGeoWords * sizeof(float4);
    if ((nodes && node_cap < size_t(c->nnodes)) || (tris && tri_cap < size_t(c->ntri)))
        return c->fail(RM_ERR_INVALID, "rm_read_tree: output arrays too small");
    if (int rc = check(c, hipStreamSynchronize(c->stream), "readback")) return rc;
    if (nodes && nb && hipMemcpy(nodes, c->nodes.ptr, nb, hipMemcpyDeviceToHost) != hipSuccess)
        return c->fail(RM_ERR_DEVICE, "readback failed");
    if (tris && tb && hipMemcpy(tris, c->geo.ptr, tb, hipMemcpyDeviceToHost) != hipSuccess)
        return c->fail(RM_ERR_DEVICE, "readback failed");
    return RM_OK;
}

int rm_trace_device(rm_ctx* c, const float* origins, const float* dirs, const float* tmax, size_t n, int any,
                    int32_t* tri, float* tuv) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    if (n == 0) return RM_OK;
    if (n > kTraceDeviceMaxRays)
        return c->fail(RM_ERR_INVALID, "rm_trace_device: at most 2^30 rays per call");
    // Every buffer must be device memory of this device holding the bytes the
    // kernel touches (a host pointer would fault the GPU instead of failing).
    struct Arg { const void* p; size_t bytes; const char* name; };
    const Arg args[] = {{origins, n * 12, "origins"}, {dirs, n * 12, "dirs"}, {tmax, n * 4, "tmax"},
                        {tri, n * 4, "tri"}, {tuv, n * 16, "tuv"}};
    for (const Arg& a : args) {
        if (!a.p) {
            if (a.p == tmax) continue;
            return c->fail(RM_ERR_INVALID, std::string("rm_trace_device: ") + a.name + " is NULL");
        }
        if (reinterpret_cast<uintptr_t>(a.p) % (a.p == tuv ? 16 : 4) != 0)
            return c->fail(RM_ERR_INVALID, std::string("rm_trace_device: ") + a.name + " is misaligned");
        hipPointerAttribute_t at{};
        void* base = nullptr;
        size_t size = 0;
        if (hipPointerGetAttributes(&at, a.p) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != c->device ||
            hipMemGetAddressRange(&base, &size, const_cast<void*>(a.p)) != hipSuccess ||
            static_cast<const char*>(a.p) + a.bytes > static_cast<const char*>(base) + size) {
            (void)hipGetLastError();
            return c->fail(RM_ERR_INVALID, std::string("rm_trace_device: ") + a.name +
                                               " is not device memory of this GPU holding the call's rays");
        }
    }
    hipStream_t s = c->stream;
    launch_zero_u64(c->scratch_dev, kNumCounters, s);
    const DevScene sc = c->dev_scene();
    launch_trace(c->stack_need, sc, FrameParams{}, int(n), origins, dirs, tmax, any, tri, reinterpret_cast<float4*>(tuv),
                 c->scratch_dev, s);
    if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
    if (int rc = check(c, hipMemcpyAsync(c->scratch_host, c->scratch_dev, kNumCounters * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, s), "counters"))
        return rc;
    if (int rc = check(c, hipStreamSynchronize(s), "trace")) return rc;
    if (c->scratch_host[kCntOverflow]) return c->fail(RM_ERR_DEVICE, "traversal stack overflow");
    rm_stats st{};
    collect(st, c->scratch_host);
    st.bvh_nodes = c->nnodes;
    st.stack_depth = c->stack_need;
    c->stats = st;
    c->err.clear();
    return RM_OK;
}

int rm_render_adjoint(rm_ctx* c, const rm_frame* f, uint64_t seed, const float* d_hdr, const int32_t* tri_group,
                      const float* group_emission, int ngroup, double* grad) {
    if (!c || !f || !d_hdr || !tri_group || !group_emission || !grad) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (f->width <= 0 || f->height <= 0 || f->spp <= 0 || f->bounces < 0 ||
        size_t(f->width) * size_t(f->height) > 100000000)
        return c->fail(RM_ERR_INVALID, "Invalid render dimensions or sample count");
    if (ngroup < 1 || ngroup > kMaxGroups)
        return c->fail(RM_ERR_INVALID, "adjoint: ngroup must be in [1, " + std::to_string(kMaxGroups) + "]");
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    const FrameParams P = frame_params(*f, seed);
    if (uses_restir(c, P))
        return c->fail(RM_ERR_INVALID, "adjoint: the frame takes the ReSTIR path; differentiate the path-tracing "
                                       "frame (restir_spatial = 0)");
    if (P.restir_candidates > 1)
        return c->fail(RM_ERR_INVALID, "adjoint: RIS light sampling is not differentiated (restir_candidates <= 1)");
    if (P.adapt_min > 0 && P.adapt_min <= P.spp)
        return c->fail(RM_ERR_INVALID, "adjoint: adaptive sampling must be off (adapt_min <= 0 or > spp)");
    for (size_t i = 0; i < size_t(c->ntri); ++i)
        if (tri_group[i] >= ngroup) return c->fail(RM_ERR_INVALID, "adjoint: triangle group out of range");
    const size_t px = size_t(f->width) * size_t(f->height);
    std::vector<int32_t> lg(c->light_tri.size());
    for (size_t k = 0; k < lg.size(); ++k) lg[k] = tri_group[c->light_tri[k]];
    const int lanes = megakernel_lanes(P);
    hipStream_t st = c->stream;
    if (!c->adj_img.ensure(px * 3 * sizeof(float)) || !c->adj_tri_group.ensure(size_t(c->ntri) * 4) ||
        !c->adj_light_group.ensure(lg.size() * 4) || !c->adj_base.ensure(size_t(ngroup) * 3 * sizeof(float)) ||
        !c->adj_grad.ensure(size_t(ngroup) * 5 * sizeof(double)) || !c->adj_bad.ensure(sizeof(unsigned long long)) ||
        !c->adj_primal.ensure(px * size_t(lanes) * sizeof(float4)) ||
        !c->adj_record.ensure(px * size_t(lanes) * sizeof(uint4)))
        return c->fail(RM_ERR_ALLOC, "HIP adjoint buffer allocation failed");
    unsigned long long* cnt = c->counters.as<unsigned long long>();
    auto h2d = [&](DevBuf& b, const void* src, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    if (!h2d(c->adj_img, d_hdr, px * 3 * sizeof(float)) || !h2d(c->adj_tri_group, tri_group, size_t(c->ntri) * 4) ||
        !h2d(c->adj_light_group, lg.data(), lg.size() * 4) ||
        !h2d(c->adj_base, group_emission, size_t(ngroup) * 3 * sizeof(float)) ||
        hipMemsetAsync(c->adj_grad.ptr, 0, size_t(ngroup) * 5 * sizeof(double), st) != hipSuccess ||
        hipMemsetAsync(c->adj_bad.ptr, 0, sizeof(unsigned long long), st) != hipSuccess ||
        hipMemsetAsync(cnt, 0, kNumCounters * sizeof(unsigned long long), st) != hipSuccess)
        return c->fail(RM_ERR_DEVICE, "adjoint: upload failed");
    AdjointArgs A;
    A.d_img = c->adj_img.as<float>();
    A.tri_group = c->adj_tri_group.as<int32_t>();
    A.light_group = c->adj_light_group.as<int32_t>();
    A.group_base = c->adj_base.as<float>();
    A.ngroup = ngroup;
    A.grad = c->adj_grad.as<double>();
    A.mismatch = c->adj_bad.as<unsigned long long>();
    A.primal = c->adj_primal.as<float4>();
    A.record = c->adj_record.as<uint4>();
    const DevScene sc = c->dev_scene();
    if (c->profiling) (void)hipEventRecord(c->ev[0], st);
    launch_adjoint(c->stack_need, lanes, sc, P, A, cnt, st);
    if (c->profiling) (void)hipEventRecord(c->ev[1], st);
    if (int rc = check(c, hipGetLastError(), "kernel launch")) return rc;
    unsigned long long host_cnt[kNumCounters], bad = 0;
    if (int rc = check(c, hipMemcpyAsync(grad, A.grad, size_t(ngroup) * 5 * sizeof(double), hipMemcpyDeviceToHost, st),
                       "readback"))
        return rc;
    if (int rc = check(c, hipMemcpyAsync(&bad, A.mismatch, sizeof(bad), hipMemcpyDeviceToHost, st), "readback"))
        return rc;
    if (int rc = check(c, hipMemcpyAsync(host_cnt, cnt, sizeof(host_cnt), hipMemcpyDeviceToHost, st), "counters"))
        return rc;
    if (int rc = check(c, hipStreamSynchronize(st), "adjoint")) return rc;
    if (host_cnt[kCntOverflow]) return c->fail(RM_ERR_DEVICE, "traversal stack overflow");
    if (bad) return c->fail(RM_ERR_DEVICE, "adjoint: " + std::to_string(bad) + " samples replayed differently");
    rm_stats sts{};
    collect(sts, host_cnt);
    sts.path = 0;
    sts.lanes = lanes;
    sts.bvh_nodes = c->nnodes;
    sts.stack_depth = c->stack_need;
    if (c->profiling && hipEventElapsedTime(&sts.ms[7], c->ev[0], c->ev[1]) != hipSuccess) sts.ms[7] = 0.0f;
    c->stats = sts;
    c->err.clear();
    return RM_OK;
}

int rm_read_restir(rm_ctx* c, rm_gbuffer* gbuf, rm_reservoir* rinit, rm_reservoir* rmerged, float* shade) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!c->hist_valid) return c->fail(RM_ERR_INVALID, "no ReSTIR frame rendered yet");
    // hipMemcpy runs on the null stream, which does not wait for the context's stream.
    if (int rc = check(c, hipStreamSynchronize(c->stream), "readback")) return rc;
    const size_t px = size_t(c->hist_w) * size_t(c->hist_h);
    std::vector<float4> a(px), b(px), d(px), e(px), g(px);
    auto get = [&](const DevBuf& buf, std::vector<float4>& dst) {
        return hipMemcpy(dst.data(), buf.ptr, px * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess;
    };
    // After a frame the G-buffer and merged reservoirs live in the history slots.
    if (gbuf) {
        const GBufSet& s = c->hist_gbuf;
        if (!get(s.p[0], a) || !get(s.p[1], b) || !get(s.p[2], d) || !get(s.p[3], e) || !get(s.p[4], g))
            return c->fail(RM_ERR_DEVICE, "readback failed");
        for (size_t i = 0; i < px; ++i) {
            rm_gbuffer& o = gbuf[i];
            o.pos = {a[i].x, a[i].y, a[i].z};
            std::memcpy(&o.kind, &a[i].w, 4);
            o.normal = {b[i].x, b[i].y, b[i].z};
            o.rough = b[i].w;
            o.albedo = {d[i].x, d[i].y, d[i].z};
            o.metallic = d[i].w;
            o.wo = {e[i].x, e[i].y, e[i].z};
            o.ior = e[i].w;
            o.emission = {g[i].x, g[i].y, g[i].z};
        }
    }
    auto res = [&](const ResSet& s, rm_reservoir* out) {
        if (!get(s.r[0], a) || !get(s.r[1], b) || !get(s.r[2], d) || !get(s.r[3], e)) return false;
        for (size_t i = 0; i < px; ++i) {
            rm_reservoir& o = out[i];
            o.light_pos = {a[i].x, a[i].y, a[i].z};
            o.p_hat = a[i].w;
            o.light_normal = {b[i].x, b[i].y, b[i].z};
            o.alpha = b[i].w;
            o.emission = {d[i].x, d[i].y, d[i].z};
            o.sum_ws = d[i].w;
            o.sum_wu = e[i].x;
            std::memcpy(&o.m, &e[i].y, 4);
            std::memcpy(&o.valid, &e[i].z, 4);
        }
        return true;
    };
    if (rinit && !res(c->rinit, rinit)) return c->fail(RM_ERR_DEVICE, "readback failed");
    if (rmerged && !res(c->hist_res, rmerged)) return c->fail(RM_ERR_DEVICE, "readback failed");
    if (shade) {
        if (!get(c->shade, a)) return c->fail(RM_ERR_DEVICE, "readback failed");
        for (size_t i = 0; i < px; ++i) {
            shade[3 * i] = a[i].x;
            shade[3 * i + 1] = a[i].y;
            shade[3 * i + 2] = a[i].z;
        }
    }
    return RM_OK;
}

int rm_restir_merge(rm_ctx* c, const rm_frame* f, uint64_t seed, const rm_gbuffer* gbuf, const rm_reservoir* rinit,
                    const rm_gbuffer* hist_gbuf, const rm_reservoir* hist_res, const rm_camera* prev_camera,
                    rm_reservoir* rmerged) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!f || !gbuf || !rinit || !rmerged || f->width <= 0 || f->height <= 0 ||
        size_t(f->width) * size_t(f->height) > 100000000)
        return c->fail(RM_ERR_INVALID, "rm_restir_merge: invalid arguments");
    const bool hist = hist_gbuf || hist_res || prev_camera;
    if (hist && !(hist_gbuf && hist_res && prev_camera))
        return c->fail(RM_ERR_INVALID, "rm_restir_merge: history needs hist_gbuf, hist_res and prev_camera");
    if (!c->has_scene) return c->fail(RM_ERR_NO_SCENE, "no scene uploaded (rm_set_scene)");
    const size_t px = size_t(f->width) * size_t(f->height);
    FrameParams P = frame_params(*f, seed);
    P.has_history = hist ? 1 : 0;
    if (hist) P.prev = *prev_camera;
    if (!c->x_gbuf.ensure(px) || !c->x_rinit.ensure(px) || !c->x_rout.ensure(px) ||
        (hist && (!c->x_hgbuf.ensure(px) || !c->x_hres.ensure(px))))
        return c->fail(RM_ERR_ALLOC, "HIP ReSTIR buffer allocation failed");
    std::vector<float4> g[5], r[4], hg[5], hr[4];
    gbuf_to_soa(gbuf, px, g);
    res_to_soa(rinit, px, r);
    if (hist) {
        gbuf_to_soa(hist_gbuf, px, hg);
        res_to_soa(hist_res, px, hr);
    }
    if (!upload_soa(c, c->x_gbuf.p, g) || !upload_soa(c, c->x_rinit.r, r) ||
        (hist && (!upload_soa(c, c->x_hgbuf.p, hg) || !upload_soa(c, c->x_hres.r, hr))))
        return c->fail(RM_ERR_DEVICE, "rm_restir_merge: upload failed");
    unsigned long long* cnt = c->counters.as<unsigned long long>();
    if (int rc = check(c, hipMemsetAsync(cnt, 0, kNumCounters * sizeof(unsigned long long), c->stream), "counters"))
        return rc;
    // Without history the kernel never reads the history arrays (has_history 0).
    const GBufferSoA hgs = hist ? c->x_hgbuf.soa() : c->x_gbuf.soa();
    const ReservoirSoA hrs = hist ? c->x_hres.soa() : c->x_rinit.soa();
    launch_spatial(c->stack_need, c->dev_scene(), P, c->x_gbuf.soa(), c->x_rinit.soa(), c->x_rout.soa(), hgs, hrs,
                   cnt, c->stream);
    if (int rc = check(c, hipGetLastError(), "restir_spatial")) return rc;
    std::vector<float4> o[4];
    for (int k = 0; k < 4; ++k) {
        o[k].resize(px);
        if (int rc = check(c, hipMemcpyAsync(o[k].data(), c->x_rout.r[k].ptr, px * sizeof(float4),
                                             hipMemcpyDeviceToHost, c->stream), "readback"))
            return rc;
    }
    unsigned long long host_cnt[kNumCounters];
    if (int rc = check(c, hipMemcpyAsync(host_cnt, cnt, sizeof(host_cnt), hipMemcpyDeviceToHost, c->stream), "counters"))
        return rc;
    if (int rc = check(c, hipStreamSynchronize(c->stream), "restir_spatial")) return rc;
    if (host_cnt[kCntOverflow]) return c->fail(RM_ERR_DEVICE, "traversal stack overflow");
    res_from_soa(o, px, rmerged);
    c->err.clear();
    return RM_OK;
}

int rm_denoise_image(rm_ctx* c, const rm_frame* f, const float* color, const rm_gbuffer* gbuf, const float* hist_color,
                     const rm_gbuffer* hist_gbuf, const rm_camera* prev_camera, float* out) {
    if (!c) return RM_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (int rc = check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
    if (!f || !color || !gbuf || !out || f->width <= 0 || f->height <= 0 ||
        size_t(f->width) * size_t(f->height) > 100000000 || f->denoise_iterations < 0 || f->denoise_iterations > 30)
        return c->fail(RM_ERR_INVALID, "rm_denoise_image: invalid arguments");
    const bool hist = hist_color || hist_gbuf || prev_camera;
    if (hist && !(hist_color && hist_gbuf && prev_camera))
        return c->fail(RM_ERR_INVALID, "rm_denoise_image: history needs hist_color, hist_gbuf and prev_camera");
    const size_t px = size_t(f->width) * size_t(f->height);
    FrameParams P = frame_params(*f, 0);
    P.has_history = hist ? 1 : 0;
    if (hist) P.prev = *prev_camera;
    if (!c->x_gbuf.ensure(px) || !c->x_col.ensure(px * sizeof(float4)) || !c->x_a.ensure(px * sizeof(float4)) ||
        !c->x_b.ensure(px * sizeof(float4)) ||
        (hist && (!c->x_hgbuf.ensure(px) || !c->x_hcol.ensure(px * sizeof(float4)))))
        return c->fail(RM_ERR_ALLOC, "HIP denoiser buffer allocation failed");
    std::vector<float4> g[5], hg[5];
    gbuf_to_soa(gbuf, px, g);
    const std::vector<float4> col = rgb_to_float4(color, px);
    std::vector<float4> hcol;
    if (hist) {
        gbuf_to_soa(hist_gbuf, px, hg);
        hcol = rgb_to_float4(hist_color, px);
    }
    if (!upload_soa(c, c->x_gbuf.p, g) || !upload(c, c->x_col, col) ||
        (hist && (!upload_soa(c, c->x_hgbuf.p, hg) || !upload(c, c->x_hcol, hcol))))
        return c->fail(RM_ERR_DEVICE, "rm_denoise_image: upload failed");
    float4* a = c->x_a.as<float4>();
    float4* b = c->x_b.as<float4>();
    const GBufferSoA gs = c->x_gbuf.soa();
    launch_temporal(P, gs, hist ? c->x_hgbuf.soa() : gs, c->x_col.as<float4>(),
                    hist ? c->x_hcol.as<float4>() : c->x_col.as<float4>(), a, c->stream);
    for (int s = 0; s < f->denoise_iterations; ++s) {
        launch_atrous(P, gs, a, b, s, c->stream);
        std::swap(a, b);
    }
    if (int rc = check(c, hipGetLastError(), "denoise")) return rc;
    std::vector<float4> res(px);
    if (int rc = check(c, hipMemcpyAsync(res.data(), a, px * sizeof(float4), hipMemcpyDeviceToHost, c->stream),
                       "readback"))
        return rc;
    if (int rc = check(c, hipStreamSynchronize(c->stream), "denoise")) return rc;
    for (size_t i = 0; i < px; ++i) {
        out[3 * i] = res[i].x;
        out[3 * i + 1] = res[i].y;
        out[3 * i + 2] = res[i].z;
    }
    c->err.clear();
    return RM_OK;
}

}  // extern "C"
