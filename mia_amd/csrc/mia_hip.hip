// C-ABI implementation (include/mia_hip.h): context, device buffers, spline
// axis tables and the launch sequence of one cost + gradient evaluation.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <dlfcn.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdlib>
#include <map>
#include <unordered_map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/mia_hip.h"
#include "common.hh"
#include "kernels.hh"
#include "optimizer.hh"

using namespace mia_hip;

namespace mia_hip_rt {
namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

const char *last_error()
{
    return g_last_error.c_str();
}

namespace {
std::mutex g_stream_mu;
std::vector<std::pair<int, hipStream_t>> g_idle_streams; // (device, stream)
constexpr int kIdleStreamsPerDevice = 8;
} // namespace

hipError_t stream_acquire(int device, hipStream_t *s)
{
    {
        std::lock_guard<std::mutex> lock(g_stream_mu);
        for (size_t k = g_idle_streams.size(); k-- > 0;)
            if (g_idle_streams[k].first == device) {
                *s = g_idle_streams[k].second;
                g_idle_streams.erase(g_idle_streams.begin() + (long)k);
                return hipSuccess;
            }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

void stream_release(int device, hipStream_t s)
{
    if (!s)
        return;
    (void)hipSetDevice(device);
    if (hipStreamSynchronize(s) == hipSuccess) {
        std::lock_guard<std::mutex> lock(g_stream_mu);
        int kept = 0;
        for (auto &e : g_idle_streams)
            kept += e.first == device;
        if (kept < kIdleStreamsPerDevice) {
            g_idle_streams.emplace_back(device, s);
            return;
        }
    }
    (void)hipStreamDestroy(s);
}

// Streams of mia_hip_optimize_many, per device: created together with the
// device's first context (the first streams of the library on that device),
// so that HIP gives them hardware queues of their own, and kept until the
// process ends.
constexpr int kJointStreams = 3;
namespace {
std::mutex g_joint_mu;
std::vector<std::pair<int, std::array<hipStream_t, kJointStreams>>> g_joint;
} // namespace

hipError_t joint_streams(int device, std::array<hipStream_t, kJointStreams> *out)
{
    std::lock_guard<std::mutex> lock(g_joint_mu);
    for (auto &j : g_joint)
        if (j.first == device) {
            if (out)
                *out = j.second;
            return hipSuccess;
        }
    std::array<hipStream_t, kJointStreams> s{};
    for (int k = 0; k < kJointStreams; ++k) {
        hipError_t e = hipStreamCreateWithFlags(&s[k], hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (int j = 0; j < k; ++j)
                (void)hipStreamDestroy(s[j]);
            return e;
        }
    }
    g_joint.emplace_back(device, s);
    if (out)
        *out = s;
    return hipSuccess;
}

} // namespace mia_hip_rt

namespace {

using mia_hip_rt::fail;

enum Stage { ST_WARP = 0, ST_STAGE1, ST_FINALIZE, ST_REDUCE, ST_COUNT };
const char *kStageNames[ST_COUNT] = {"warp", "ssd_force_adjoint1", "finalize", "adjoint_reduce"};

// Spline axis table (entries from spline_axis_entry) and its band table:
// cells[c] = first sample whose start index is >= c.
struct AxisTable {
    std::vector<float> w; // 4 per sample
    std::vector<int> start;
    std::vector<int> cells; // csize + 1
    int nbands = 0;
};

// Sizes an axis of the kernels accepts, checked without building the table:
// the number of knot bands and the most rows (samples) in one band.
struct AxisCheck {
    int nbands = 0, band_rows = 0;
};

int check_axis(int n, int cs, const char *axis, AxisCheck &out)
{
    if (n < 2)
        return fail(MIA_HIP_ERR_ARG, std::string("axis ") + axis + ": image size must be >= 2");
    if (cs < 4)
        return fail(MIA_HIP_ERR_ARG, std::string("axis ") + axis + ": coefficient grid must be >= 4 (2 + enlarge)");
    if (cs - 3 > n - 1)
        return fail(MIA_HIP_ERR_LIMIT, std::string("axis ") + axis +
                                           ": coefficient grid denser than the image (scale > 1) is not supported");
    float w[4];
    int prev = spline_axis_entry(0, n, cs, w), run = 1, rows = 1, last_but_one = prev;
    if (prev != 0)
        return fail(MIA_HIP_ERR_ARG, std::string("axis ") + axis + ": unexpected first start index");
    for (int i = 1; i < n; ++i) {
        const int s = spline_axis_entry(i, n, cs, w);
        const int dlt = s - prev;
        if (dlt < 0 || dlt > 1)
            return fail(MIA_HIP_ERR_LIMIT, std::string("axis ") + axis + ": non-unit knot step");
        run = dlt ? 1 : run + 1;
        rows = std::max(rows, run);
        last_but_one = prev;
        prev = s;
    }
    if (prev + 3 > cs || (n > 1 && last_but_one + 3 > cs - 1))
        return fail(MIA_HIP_ERR_ARG, std::string("axis ") + axis + ": knot range exceeds the coefficient grid");
    out.nbands = prev + 1;
    out.band_rows = rows;
    return MIA_HIP_OK;
}

int make_axis_table(int n, int cs, AxisTable &t, const char *axis)
{
    AxisCheck ck;
    int st = check_axis(n, cs, axis, ck);
    if (st)
        return st;
    t.w.assign(4 * (size_t)n, 0.0f);
    t.start.assign(n, 0);
    for (int i = 0; i < n; ++i)
        t.start[i] = spline_axis_entry(i, n, cs, &t.w[4 * (size_t)i]);
    t.cells.assign(cs + 1, n);
    int c = 0;
    for (int i = 0; i < n; ++i)
        while (c <= t.start[i]) {
            t.cells[c] = i;
            ++c;
        }
    t.nbands = t.start[n - 1] + 1;
    return MIA_HIP_OK;
}

long long align64(long long v)
{
    return (v + 63) & ~63LL;
}

} // namespace

struct mia_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int dim = 0, n_inst = 0, degree = 3;
    int norm = 0, grad_mode = MIA_HIP_GRAD_REFERENCE;
    float thr = 0.0f;
    double weight = 1.0;

    std::vector<InstDesc> inst;
    std::vector<int> s1_prefix, red_prefix;
    // 3D reference gradient: stage-1 blocks per (z knot cell, y band), reduce
    // blocks of kBlock gradient components
    std::vector<int> s1z_prefix, redz_prefix;
    // warp launches: instances grouped by tile variant (one launch per group)
    struct WarpGroup {
        int variant = 0;           // 3D tile variant; unused in 2D
        std::vector<int> ids;      // instances of the group, ascending
        std::vector<int> prefix;   // tile prefix over ids
        int ids_off = 0, prefix_off = 0; // offsets into d_wgroups
        int tiles2d_off = -1;            // 2D: offset of the tile table in d_wgroups (multiple of 4)
    };
    std::vector<WarpGroup> wgroups;
    int s1tab_off = -1; // 2D: offset in d_wgroups of the stage-1 block table (multiple of 4)
    int redtab_off = -1; // 2D: offset in d_wgroups of the reduce block table (multiple of 2)
    std::vector<int> inst_variant, inst_group, inst_slot;
    int force_variant = -1; // -1: tuned per size bucket
    int force_s1 = -1;      // stage-1 layout of 3D reference gradients: -1 tuned, 0 per slice, 1 z cell
    std::vector<char> images_set;
    long long n_img = 0, n_par = 0, n_part = 0, max_vox = 0;
    int max_nx = 0, max_cx = 0, max_ncx = 0, n_zcell = 0;
    int w2_win = 0; // 2D warp: largest coefficient window of a tile (float2)

    InstDesc *d_inst = nullptr;
    int *d_wgroups = nullptr, *d_s1_prefix = nullptr, *d_red_prefix = nullptr, *d_s1z_prefix = nullptr,
        *d_redz_prefix = nullptr;
    float *d_M = nullptr, *d_R = nullptr, *d_W = nullptr, *d_P = nullptr;
    float2 *d_clamp = nullptr, *d_mm = nullptr;
    float4 *d_tw = nullptr;
    int *d_ts = nullptr, *d_cells = nullptr;
    double *d_ssd = nullptr, *d_gscale = nullptr, *d_params = nullptr, *d_grad = nullptr, *d_costs = nullptr;
    double *d_scratch = nullptr;
    unsigned *d_cnt = nullptr;

    // pinned host buffers, kept across set-ups and grown when a set-up needs
    // more (capacities in elements): the graph's parameter, gradient and cost
    // staging, the optimizer's done flags, and the staging of
    // small uploads (a running offset, reset only once the stream is idle)
    double *h_params = nullptr, *h_grad = nullptr, *h_costs = nullptr;
    size_t cap_params = 0, cap_grad = 0, cap_costs = 0, cap_done = 0;
    char *h_up = nullptr;
    size_t cap_up = 0, up_off = 0;
    // on-device optimizer state (allocated on first use, sized for `opt_memory`)
    int opt_memory = 0;
    double *d_oX = nullptr, *d_oG = nullptr, *d_oD = nullptr, *d_oBX = nullptr, *d_oS = nullptr, *d_oY = nullptr,
           *d_obest = nullptr, *d_oSY = nullptr, *d_oYY = nullptr, *d_oPART = nullptr, *d_oPART2 = nullptr;
    int *d_oseg = nullptr;
    int opt_segments = 0; // segments of the optimizer's passes over all instances
    int opt_launch = MIA_HIP_LAUNCH_STREAM; // how mia_hip_optimize queues its chunks
    // done flags, written by the optimizer's kernels straight into this
    // pinned host memory (no copy per chunk; a flag only goes from 0 to 1)
    int *h_done = nullptr;
    OptState *d_ost = nullptr, *d_ost2 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;

    // spline penalties evaluated with every evaluation, per instance (none: null)
    std::vector<mia_hip_divcurl *> penalty;
    std::vector<double> penalty_weight;
    // their device inputs of the finalize and reduce (allocated with the
    // first attached penalty) and the second stream they are evaluated on,
    // forked from and joined back into the context stream in every evaluation
    int *d_pen_on = nullptr;
    double *d_pen_w = nullptr, *d_pen_value = nullptr, *d_pen_grad = nullptr;
    hipStream_t pen_stream = nullptr;
    hipEvent_t pen_fork = nullptr, pen_join = nullptr;

    // device blocks this context allocated (bytes), and those of them free
    // for reuse: a set-up re-allocates from the blocks the previous set-up
    // released, and returns the rest to HIP (hipMalloc and hipFree cost
    // tens to hundreds of microseconds each)
    std::unordered_map<void *, size_t> blocks;
    std::multimap<size_t, void *> idle;

    bool profile = false;
    struct Ev {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Ev> pending;
    std::vector<hipEvent_t> pool;
    double prof_ms[ST_COUNT] = {0};
    int prof_n[ST_COUNT] = {0};
};

namespace {

// Return a block of this context to its idle pool.
void pool_release(mia_hip_ctx *c, void *p)
{
    if (!p)
        return;
    auto it = c->blocks.find(p);
    if (it == c->blocks.end())
        (void)hipFree(p);
    else
        c->idle.emplace(it->second, p);
}

// hipFree the idle blocks.
void pool_trim(mia_hip_ctx *c)
{
    for (auto &kv : c->idle) {
        c->blocks.erase(kv.second);
        (void)hipFree(kv.second);
    }
    c->idle.clear();
}

void free_buffers(mia_hip_ctx *c)
{
    void *ptrs[] = {c->d_inst, c->d_wgroups, c->d_s1_prefix, c->d_red_prefix, c->d_s1z_prefix, c->d_redz_prefix,
                    c->d_M, c->d_R, c->d_W,
                    c->d_P, c->d_clamp, c->d_mm, c->d_tw, c->d_ts, c->d_cells, c->d_ssd, c->d_gscale,
                    c->d_params, c->d_grad, c->d_costs, c->d_scratch, c->d_cnt, c->d_oX, c->d_oG, c->d_oD,
                    c->d_oBX, c->d_oS, c->d_oY, c->d_obest, c->d_ost, c->d_oSY,
                    c->d_oYY, c->d_oPART, c->d_oPART2, c->d_oseg, c->d_ost2, c->d_pen_on, c->d_pen_w,
                    c->d_pen_value, c->d_pen_grad};
    // blocks the previous set-up released and this one did not reuse go
    // back to HIP; this set-up's join the pool (the stream is idle here)
    pool_trim(c);
    for (void *p : ptrs)
        pool_release(c, p);
    c->d_oX = c->d_oG = c->d_oD = c->d_oBX = c->d_oS = c->d_oY = c->d_obest = nullptr;
    c->d_oSY = c->d_oYY = c->d_oPART = c->d_oPART2 = nullptr;
    c->d_oseg = nullptr;
    c->d_ost2 = nullptr;
    c->opt_segments = 0;
    c->d_ost = nullptr;
    c->opt_memory = 0;
    if (c->gexec)
        (void)hipGraphExecDestroy(c->gexec);
    if (c->graph)
        (void)hipGraphDestroy(c->graph);
    c->d_inst = nullptr;
    c->d_wgroups = c->d_s1_prefix = c->d_red_prefix = c->d_s1z_prefix = c->d_redz_prefix = nullptr;
    c->d_M = c->d_R = c->d_W = c->d_P = nullptr;
    c->d_clamp = c->d_mm = nullptr;
    c->d_tw = nullptr;
    c->d_ts = c->d_cells = nullptr;
    c->d_ssd = c->d_gscale = c->d_params = c->d_grad = c->d_costs = c->d_scratch = nullptr;
    c->d_cnt = nullptr;
    c->d_pen_on = nullptr;
    c->d_pen_w = c->d_pen_value = c->d_pen_grad = nullptr;
    c->gexec = nullptr;
    c->graph = nullptr;
}

// A pinned host buffer of at least `count` elements, kept from earlier
// calls when large enough (callers make sure no queued copy still uses it).
template <typename T> int host_reserve(T **p, size_t *cap, size_t count)
{
    count = std::max<size_t>(count, 1);
    if (*p && *cap >= count)
        return MIA_HIP_OK;
    if (*p)
        (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    HIP_TRY(hipHostMalloc((void **)p, sizeof(T) * count, hipHostMallocDefault));
    *cap = count;
    return MIA_HIP_OK;
}

// Queue an upload of host memory through the context's pinned staging (a
// copy from pageable memory takes the runtime's slower staged path): the
// bytes are copied into the staging buffer at its running offset and sent
// from there on the context's stream.  When the buffer is full the stream is
// drained first, so that no queued copy still reads what is overwritten.
int stage_upload(mia_hip_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0)
        return MIA_HIP_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (c->up_off + need > c->cap_up) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->up_off = 0;
        if (need > c->cap_up) {
            int st = host_reserve(&c->h_up, &c->cap_up, std::max<size_t>(need, std::max<size_t>(2 * c->cap_up, 1 << 20)));
            if (st)
                return st;
        }
    }
    std::memcpy(c->h_up + c->up_off, src, bytes);
    HIP_TRY(hipMemcpyAsync(dst, c->h_up + c->up_off, bytes, hipMemcpyHostToDevice, c->stream));
    c->up_off += need;
    return MIA_HIP_OK;
}

// Read device memory back through the pinned staging, synchronously (the
// stream is drained, so the staging offset starts again from zero).
int stage_download(mia_hip_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0)
        return MIA_HIP_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->up_off = 0;
    if (bytes > c->cap_up) {
        int st = host_reserve(&c->h_up, &c->cap_up, std::max<size_t>(bytes, std::max<size_t>(2 * c->cap_up, 1 << 20)));
        if (st)
            return st;
    }
    HIP_TRY(hipMemcpyAsync(c->h_up, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(dst, c->h_up, bytes);
    return MIA_HIP_OK;
}

// Tuned choices per (kind, interpolation degree, size bucket = ceil(log2(voxels
// or pixels))): the 3D warp tile, the 2D warp tile and the stage-1 layout of
// 3D reference gradients (0 per slice, 1 per z knot cell).  One JSON table
// keyed by the GPU architecture, written by mia_hip_autotune.  Source:
// $MIA_HIP_TUNING, else <library dir>/../tuning/<arch>.json; without an
// entry for a kind and degree, the kind's default.
constexpr int kTuneKinds = 3; // MIA_HIP_TUNE_WARP3D, _WARP2D, _STAGE1
const char *const kTuneKindNames[kTuneKinds] = {"warp3d", "warp2d", "stage1_3d"};
constexpr int kTuneChoices[kTuneKinds] = {kW3Variants, kW2Variants, 2};
constexpr int kTuneDefault[kTuneKinds] = {0, 0, 1};
constexpr int kMaxTuneChoices = 6;
static_assert(kW3Variants <= kMaxTuneChoices && kW2Variants <= kMaxTuneChoices, "variant tables");

struct Tuning {
    std::mutex m;
    bool loaded = false;
    std::string arch, source;
    std::map<std::tuple<int, int, int>, int> table; // (kind, degree, bucket) -> choice
    // (kind, degree, bucket) -> the tuned case's time of every choice (us), where the table has them
    std::map<std::tuple<int, int, int>, std::vector<double>> us;
};

Tuning &tuning()
{
    static Tuning t;
    return t;
}

std::string device_arch(int device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return "";
    std::string a(prop.gcnArchName);
    return a.substr(0, a.find(':'));
}

std::string default_tuning_path(const std::string &arch)
{
    if (const char *e = std::getenv("MIA_HIP_TUNING"))
        return e;
    Dl_info info;
    if (dladdr((const void *)&default_tuning_path, &info) && info.dli_fname) {
        std::string lib(info.dli_fname);
        const size_t k = lib.rfind('/');
        const std::string dir = k == std::string::npos ? "." : lib.substr(0, k);
        return dir + "/../tuning/" + arch + ".json";
    }
    return "";
}

bool read_text(const std::string &path, std::string &text)
{
    std::FILE *f = std::fopen(path.c_str(), "r");
    if (!f)
        return false;
    text.clear();
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0)
        text.append(buf, n);
    std::fclose(f);
    return true;
}

// the string value of "key" in text[b, e), or "" without one
std::string json_string(const std::string &text, size_t b, size_t e, const char *key)
{
    const std::string k = std::string("\"") + key + "\"";
    const size_t a = text.find(k, b);
    if (a == std::string::npos || a >= e)
        return "";
    const size_t q1 = text.find('"', text.find(':', a + k.size()) + 1);
    const size_t q2 = q1 == std::string::npos ? q1 : text.find('"', q1 + 1);
    if (q2 == std::string::npos || q2 >= e)
        return "";
    return text.substr(q1 + 1, q2 - q1 - 1);
}

// the integer value of "key" in text[b, e); false without one
bool json_int(const std::string &text, size_t b, size_t e, const char *key, int &v)
{
    const std::string k = std::string("\"") + key + "\"";
    const size_t a = text.find(k, b);
    if (a == std::string::npos || a >= e)
        return false;
    const size_t c = text.find(':', a + k.size());
    if (c == std::string::npos || c >= e)
        return false;
    char *end = nullptr;
    const long x = std::strtol(text.c_str() + c + 1, &end, 10);
    if (end == text.c_str() + c + 1)
        return false;
    v = (int)x;
    return true;
}

// the array of numbers "key": [a, b, ...] in text[b, e); false without one
bool json_doubles(const std::string &text, size_t b, size_t e, const char *key, std::vector<double> &v)
{
    const std::string k = std::string("\"") + key + "\"";
    const size_t a = text.find(k, b);
    if (a == std::string::npos || a >= e)
        return false;
    const size_t o = text.find('[', a + k.size()), cl = text.find(']', a + k.size());
    if (o == std::string::npos || cl == std::string::npos || o > cl || cl >= e)
        return false;
    v.clear();
    const char *q = text.c_str() + o + 1, *end = text.c_str() + cl;
    while (q < end) {
        char *nx = nullptr;
        const double x = std::strtod(q, &nx);
        if (nx == q)
            break;
        v.push_back(x);
        q = nx;
        while (q < end && (*q == ',' || *q == ' ' || *q == '\n'))
            ++q;
    }
    return !v.empty();
}

int tune_kind(const std::string &name)
{
    if (name.empty())
        return 0; // entries without "kernel" are 3D warp tiles
    for (int k = 0; k < kTuneKinds; ++k)
        if (name == kTuneKindNames[k])
            return k;
    return -1;
}

// the entries of a table written by mia_hip_autotune: objects with "bucket",
// "degree", "variant" and (optionally) "kernel"; false if unreadable
bool parse_tuning(const std::string &path, std::string &arch, std::map<std::tuple<int, int, int>, int> &table,
                  std::map<std::tuple<int, int, int>, std::vector<double>> *us = nullptr)
{
    std::string text;
    if (!read_text(path, text))
        return false;
    arch = json_string(text, 0, text.size(), "arch");
    if (arch.empty())
        return false;
    table.clear();
    if (us)
        us->clear();
    size_t pos = 0;
    while ((pos = text.find("\"bucket\"", pos)) != std::string::npos) {
        const size_t b = text.rfind('{', pos), e = text.find('}', pos);
        ++pos;
        if (b == std::string::npos || e == std::string::npos)
            continue;
        const int kind = tune_kind(json_string(text, b, e, "kernel"));
        int d = 0, bk = 0, v = 0;
        if (kind >= 0 && json_int(text, b, e, "degree", d) && json_int(text, b, e, "bucket", bk) &&
            json_int(text, b, e, "variant", v) && v >= 0 && v < kTuneChoices[kind]) {
            table[std::make_tuple(kind, d, bk)] = v;
            std::vector<double> t;
            if (us && json_doubles(text, b, e, "us", t) && (int)t.size() == kTuneChoices[kind])
                (*us)[std::make_tuple(kind, d, bk)] = t;
        }
    }
    return true;
}

void load_tuning_locked(Tuning &t, int device, const char *path)
{
    t.loaded = true;
    t.table.clear();
    t.source.clear();
    const std::string arch = device_arch(device);
    const std::string p = path ? std::string(path) : default_tuning_path(arch);
    std::string file_arch;
    std::map<std::tuple<int, int, int>, int> table;
    std::map<std::tuple<int, int, int>, std::vector<double>> us;
    t.us.clear();
    if (!p.empty() && parse_tuning(p, file_arch, table, &us) && file_arch == arch) {
        t.table = table;
        t.us = us;
        t.source = p;
    }
    t.arch = arch;
}

int size_bucket(long long n)
{
    int bucket = 0;
    while ((1LL << bucket) < n)
        ++bucket;
    return bucket;
}

// the tuned choice of `kind` for an instance of n voxels (pixels): the entry
// of the nearest bucket of that kind and degree, else the kind's default
int tuned_choice(int device, int kind, int degree, long long n)
{
    Tuning &t = tuning();
    std::lock_guard<std::mutex> lock(t.m);
    if (!t.loaded)
        load_tuning_locked(t, device, nullptr);
    const int bucket = size_bucket(n);
    int best = kTuneDefault[kind], dist = 1 << 30;
    for (const auto &kv : t.table)
        if (std::get<0>(kv.first) == kind && std::get<1>(kv.first) == degree &&
            std::abs(std::get<2>(kv.first) - bucket) < dist) {
            dist = std::abs(std::get<2>(kv.first) - bucket);
            best = kv.second;
        }
    return best;
}

// One 2D warp variant for a whole batch.  Every variant computes the same
// image (same weights and contraction order; the tile only decides which
// block computes a pixel), so a batch whose instances fall into size buckets
// with different tuned variants can run one launch instead of one per
// variant: the variant with the least tuned time summed over the instances'
// (nearest) buckets, each bucket's time being that of a batch of like
// instances.  -1 when some bucket has no timings in the table (then each
// instance keeps its own bucket's choice).
int batch_warp2d_choice(int device, int degree, const std::vector<long long> &npix)
{
    Tuning &t = tuning();
    std::lock_guard<std::mutex> lock(t.m);
    if (!t.loaded)
        load_tuning_locked(t, device, nullptr);
    const int kind = MIA_HIP_TUNE_WARP2D;
    std::vector<double> sum(kTuneChoices[kind], 0.0);
    for (long long n : npix) {
        const int bucket = size_bucket(n);
        const std::vector<double> *best = nullptr;
        int dist = 1 << 30;
        for (const auto &kv : t.table)
            if (std::get<0>(kv.first) == kind && std::get<1>(kv.first) == degree &&
                std::abs(std::get<2>(kv.first) - bucket) < dist) {
                dist = std::abs(std::get<2>(kv.first) - bucket);
                const auto it = t.us.find(kv.first);
                best = it == t.us.end() ? nullptr : &it->second;
            }
        if (!best)
            return -1;
        for (int v = 0; v < kTuneChoices[kind]; ++v)
            sum[v] += (*best)[v];
    }
    int v = 0;
    for (int k = 1; k < kTuneChoices[kind]; ++k)
        if (sum[k] < sum[v])
            v = k;
    return v;
}

// whether an instance may run the z-cell stage 1 (one x column per thread,
// the band's rows in registers; its reduce indexes the gradient with 32-bit
// arithmetic)
bool zcell_eligible(int dim, int nx, int band_rows, long long ncoef)
{
    return dim == 3 && nx <= kBlock && band_rows + 2 <= kS1zRows && 3LL * ncoef < (1LL << 31);
}

// the coefficient window of a 2D warp tile: knots spanned by the tiles of
// one axis (t pixels each), from the axis' start indices
template <typename StartOf> int tile_span(int n, int t, StartOf start)
{
    int m = 0;
    for (int p0 = 0; p0 < n; p0 += t)
        m = std::max(m, start(std::min(p0 + t, n) - 1) + 4 - start(p0));
    return m;
}

hipEvent_t get_event(mia_hip_ctx *c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    // timing only: no system-scope fence (no cache writeback between the
    // timed kernels)
    (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    return e;
}

// Launch a kernel on stream s.  With a profiling context, a pair of timing
// events brackets the dispatch itself (hipExtLaunchKernelGGL) rather than
// separate markers queued around it.
template <typename K, typename... Args>
void launch(hipStream_t s, mia_hip_ctx *prof, int stage, K kern, dim3 grid, size_t lds, Args... args)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (prof) {
        a = get_event(prof);
        b = get_event(prof);
    }
    hipExtLaunchKernelGGL(kern, grid, dim3(kBlock), (uint32_t)lds, s, a, b, 0, args...);
    if (prof)
        prof->pending.push_back({stage, a, b});
}

// A device block of `count` elements for context c: an idle block of the
// context of at least that size and at most twice it (plus 1 MiB), else a new
// one (after returning the idle blocks to HIP if it cannot be had).
template <typename T>
int dev_alloc(mia_hip_ctx *c, T **p, size_t count)
{
    const size_t bytes = ((std::max<size_t>(count, 1) * sizeof(T)) + 255) & ~(size_t)255;
    auto it = c->idle.lower_bound(bytes);
    if (it != c->idle.end() && it->first <= 2 * bytes + (1u << 20)) {
        *p = (T *)it->second;
        c->idle.erase(it);
        return MIA_HIP_OK;
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess && !c->idle.empty()) {
        (void)hipGetLastError();
        pool_trim(c);
        e = hipMalloc(&q, bytes);
    }
    if (e != hipSuccess)
        return fail(MIA_HIP_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->blocks[q] = bytes;
    *p = (T *)q;
    return MIA_HIP_OK;
}

template <int DEG>
const void *warp3d_kernel(int variant)
{
    switch (variant) {
    case 1: return (const void *)k_warp3d<DEG, W3T1>;
    case 2: return (const void *)k_warp3d<DEG, W3T2>;
    case 3: return (const void *)k_warp3d<DEG, W3T3>;
    default: return (const void *)k_warp3d<DEG, W3T0>;
    }
}

// One warp launch: the instances of a batch that share a tile variant.
struct WarpLaunch {
    int variant = 0, n = 0, tiles = 0;
    const int *ids = nullptr, *prefix = nullptr; // device: instance ids, tile prefix
    // 2D: per tile {instance, x0, y0, first x knot, first y knot, window
    // width, window height, 0} (two int4), or null (the kernel derives them)
    const int4 *tiles2d = nullptr;
};

// Everything one cost + gradient evaluation launches with: device pointers
// and host-side counts, no containers.  Built per call from a context, or on
// the stack by the allocation-free entry point over its scratch memory.
struct EvalPlan {
    int dim = 0, degree = 3, n_inst = 0, n_zcell = 0;
    int norm = 0, grad_mode = MIA_HIP_GRAD_REFERENCE;
    float thr = 0.0f;
    double weight = 1.0;
    int max_nx = 0, max_cx = 0, max_ncx = 0, w2_win = 0;
    int s1_blocks = 0, red_blocks = 0, s1z_blocks = 0, redz_blocks = 0; // prefix totals
    int n_warp = 0;
    WarpLaunch warp[kMaxTuneChoices];
    const InstDesc *inst = nullptr;
    const int *s1_prefix = nullptr, *red_prefix = nullptr, *s1z_prefix = nullptr, *redz_prefix = nullptr;
    const int4 *s1tab2d = nullptr; // 2D stage 1: per block {instance, band, first row, end row}, or null
    const int2 *redtab2d = nullptr; // 2D reduce: per gradient block {instance, block of the instance}, or null
    const float *M = nullptr, *R = nullptr;
    float *W = nullptr, *P = nullptr;
    const float4 *tw = nullptr;
    const int *ts = nullptr, *cells = nullptr;
    const float2 *clamp = nullptr;
    double *ssd = nullptr, *gscale = nullptr;
    unsigned *cnt = nullptr;
    // penalties added by the finalize and reduce, and the event their
    // evaluation on another stream records (none: null)
    PenaltyIn pen{nullptr, nullptr, nullptr, nullptr};
    hipEvent_t pen_join = nullptr;
};

EvalPlan plan_of(const mia_hip_ctx *c)
{
    EvalPlan p;
    p.dim = c->dim;
    p.degree = c->degree;
    p.n_inst = c->n_inst;
    p.n_zcell = c->n_zcell;
    p.norm = c->norm;
    p.grad_mode = c->grad_mode;
    p.thr = c->thr;
    p.weight = c->weight;
    p.max_nx = c->max_nx;
    p.max_cx = c->max_cx;
    p.max_ncx = c->max_ncx;
    p.w2_win = c->w2_win;
    const int n = c->n_inst;
    p.s1_blocks = c->s1_prefix[n];
    p.red_blocks = c->red_prefix[n];
    p.s1z_blocks = c->s1z_prefix[n];
    p.redz_blocks = c->redz_prefix[n];
    p.n_warp = (int)c->wgroups.size();
    for (int g = 0; g < p.n_warp; ++g) {
        const auto &G = c->wgroups[g];
        p.warp[g].variant = G.variant;
        p.warp[g].n = (int)G.ids.size();
        p.warp[g].tiles = G.prefix.back();
        p.warp[g].ids = c->d_wgroups + G.ids_off;
        p.warp[g].prefix = c->d_wgroups + G.prefix_off;
        p.warp[g].tiles2d = G.tiles2d_off >= 0 ? (const int4 *)(c->d_wgroups + G.tiles2d_off) : nullptr;
    }
    p.s1tab2d = c->s1tab_off >= 0 ? (const int4 *)(c->d_wgroups + c->s1tab_off) : nullptr;
    p.redtab2d = c->redtab_off >= 0 ? (const int2 *)(c->d_wgroups + c->redtab_off) : nullptr;
    p.inst = c->d_inst;
    p.s1_prefix = c->d_s1_prefix;
    p.red_prefix = c->d_red_prefix;
    p.s1z_prefix = c->d_s1z_prefix;
    p.redz_prefix = c->d_redz_prefix;
    p.M = c->d_M;
    p.R = c->d_R;
    p.W = c->d_W;
    p.P = c->d_P;
    p.tw = c->d_tw;
    p.ts = c->d_ts;
    p.cells = c->d_cells;
    p.clamp = c->d_clamp;
    p.ssd = c->d_ssd;
    p.gscale = c->d_gscale;
    p.cnt = c->d_cnt;
    return p;
}

// the 2D warp kernel of a degree and tile variant
template <int V> constexpr auto warp2d_v(int degree)
{
    return degree == 3 ? k_warp2d<3, w2_tile_x(V), w2_tile_y(V), w2_box(V)>
                       : k_warp2d<1, w2_tile_x(V), w2_tile_y(V), w2_box(V)>;
}

decltype(&k_warp2d<3, 64, 32, true>) warp2d_kernel(int degree, int v)
{
    static_assert(kW2Variants == 6, "one case per variant");
    switch (v) {
    case 1: return warp2d_v<1>(degree);
    case 2: return warp2d_v<2>(degree);
    case 3: return warp2d_v<3>(degree);
    case 4: return warp2d_v<4>(degree);
    case 5: return warp2d_v<5>(degree);
    default: return warp2d_v<0>(degree);
    }
}

// Warp the instances of warp launch g, blocks [boff, boff + nb) of its tile
// prefix.
int launch_warp(const EvalPlan &pl, hipStream_t s, mia_hip_ctx *prof, const double *d_params, int g, int boff,
                int nb)
{
    const WarpLaunch &G = pl.warp[g];
    if (nb <= 0)
        return MIA_HIP_OK;
    if (pl.dim == 3) {
        using K = void (*)(const InstDesc *, const int *, int, const int *, int, int, const double *, const float4 *,
                           const int *, const float *, const float2 *, float *);
        auto kern = (K)(pl.degree == 3 ? warp3d_kernel<3>(G.variant) : warp3d_kernel<1>(G.variant));
        launch(s, prof, ST_WARP, kern, dim3(nb), 0, pl.inst, G.ids, G.n, G.prefix, boff, nb, d_params, pl.tw, pl.ts,
               pl.M, pl.clamp, pl.W);
    } else {
        auto kern = warp2d_kernel(pl.degree, G.variant);
        const size_t lds = (size_t)w2_region(G.variant) * sizeof(float) + (size_t)pl.w2_win * sizeof(float2);
        launch(s, prof, ST_WARP, kern, dim3(nb), lds, pl.inst, G.ids, G.n, G.prefix, G.tiles2d, boff, nb, d_params,
               pl.tw, pl.ts, pl.M, pl.clamp, pl.W);
    }
    LAUNCH_CHECK("warp");
    return MIA_HIP_OK;
}

// Enqueue warp (reference mode), stage 1, finalize (only with autothresh) and
// reduce for all instances of the plan on stream s.  No memsets: every output
// element is written.
int enqueue_plan(const EvalPlan &pl, hipStream_t s, mia_hip_ctx *prof, const double *d_params, double *d_grad,
                 double *d_costs)
{
    const int n = pl.n_inst;
    const bool exact = pl.grad_mode == MIA_HIP_GRAD_EXACT;
    if (!exact)
        for (int g = 0; g < pl.n_warp; ++g) {
            int st = launch_warp(pl, s, prof, d_params, g, 0, pl.warp[g].tiles);
            if (st)
                return st;
        }
    // 3D reference gradient: instances with zcell run stage 1 per (z knot
    // cell, y band) and the single-read reduce, the others per (slice, y
    // band); their SSD partials are kept apart (SsdIndex)
    const int n_z = pl.dim == 3 && !exact ? pl.n_zcell : 0;
    const bool any_slice = n_z < n;
    SsdIndex si;
    si.s1 = pl.s1_prefix;
    si.s1z = pl.s1z_prefix;
    si.zoff = pl.s1_blocks;
    si.zmode = pl.dim == 3 && !exact ? 1 : 0;
    const size_t lds = stage1_layout(pl.dim, pl.max_nx, pl.max_cx, pl.max_ncx, exact).bytes;
    if (any_slice) {
        const int nb = pl.s1_blocks;
        if (pl.dim == 3) {
            auto kern = pl.degree == 3 ? (exact ? k_stage1_3d<3, 1> : k_stage1_3d<3, 0>)
                                       : (exact ? k_stage1_3d<1, 1> : k_stage1_3d<1, 0>);
            launch(s, prof, ST_STAGE1, kern, dim3(nb), lds, pl.inst, n, pl.s1_prefix, nb, d_params, pl.tw, pl.ts,
                   pl.cells, pl.M, pl.clamp, (const float *)pl.W, pl.R, pl.thr, pl.P, pl.ssd, pl.cnt);
        } else {
            auto kern = pl.degree == 3 ? (exact ? k_stage1_2d<3, 1> : k_stage1_2d<3, 0>)
                                       : (exact ? k_stage1_2d<1, 1> : k_stage1_2d<1, 0>);
            launch(s, prof, ST_STAGE1, kern, dim3(nb), lds, pl.inst, n, pl.s1_prefix, pl.s1tab2d, nb, d_params,
                   pl.tw, pl.ts, pl.cells, pl.M, pl.clamp, (const float *)pl.W, pl.R, pl.thr, pl.P, pl.ssd, pl.cnt);
        }
        LAUNCH_CHECK("stage1");
    }
    if (n_z > 0) {
        const int nb = pl.s1z_blocks;
        const bool thr = pl.thr != 0.0f;
        auto kern = pl.degree == 3 ? (thr ? k_stage1z_3d<3, kS1zRows, true> : k_stage1z_3d<3, kS1zRows, false>)
                                   : (thr ? k_stage1z_3d<1, kS1zRows, true> : k_stage1z_3d<1, kS1zRows, false>);
        launch(s, prof, ST_STAGE1, kern, dim3(nb), lds, pl.inst, n, pl.s1z_prefix, nb, pl.tw, pl.ts, pl.cells,
               (const float *)pl.W, pl.R, pl.thr, pl.P, pl.ssd + si.zoff, pl.cnt + si.zoff);
        LAUNCH_CHECK("stage1");
    }
    const bool fused = pl.thr == 0.0f;
    // the penalties' value and gradient, evaluated on another stream, are
    // added by the finalize and the reduce
    if (pl.pen_join)
        HIP_TRY(hipStreamWaitEvent(s, pl.pen_join, 0));
    if (!fused) {
        launch(s, prof, ST_FINALIZE, k_finalize, dim3(n), 0, pl.inst, si, (const double *)pl.ssd,
               (const unsigned *)pl.cnt, pl.norm, pl.thr, pl.weight, d_costs, pl.gscale, pl.pen);
        LAUNCH_CHECK("finalize");
    }
    // with a fused finalize, the first reduce launch does it for every instance
    int nfin = fused ? n : 0;
    if (any_slice) {
        const int nb = pl.red_blocks + nfin;
        decltype(&k_reduce<3, true>) kern;
        if (pl.dim == 3)
            kern = fused ? k_reduce<3, true> : k_reduce<3, false>;
        else
            kern = fused ? k_reduce<2, true> : k_reduce<2, false>;
        launch(s, prof, ST_REDUCE, kern, dim3(nb), 0, pl.inst, n, pl.red_prefix, pl.dim == 2 ? pl.redtab2d : nullptr,
               (const float *)pl.P, pl.tw, pl.ts, pl.cells, si, nfin, (const double *)pl.ssd,
               (const unsigned *)pl.cnt, pl.norm, pl.thr, pl.weight, d_costs, (const double *)pl.gscale, d_grad,
               pl.pen);
        LAUNCH_CHECK("reduce");
        nfin = 0;
    }
    if (n_z > 0) {
        const int nb = pl.redz_blocks + nfin;
        auto kern = fused ? k_reducez_3d<true> : k_reducez_3d<false>;
        launch(s, prof, ST_REDUCE, kern, dim3(nb), 0, pl.inst, n, pl.redz_prefix, (const float *)pl.P, si, nfin,
               (const double *)pl.ssd, (const unsigned *)pl.cnt, pl.norm, pl.thr, pl.weight, d_costs,
               (const double *)pl.gscale, d_grad, pl.pen);
        LAUNCH_CHECK("reduce");
    }
    return MIA_HIP_OK;
}

int enqueue_eval(mia_hip_ctx *c, const double *d_params, double *d_grad, double *d_costs, bool timed)
{
    EvalPlan pl = plan_of(c);
    bool any = false;
    for (auto *p : c->penalty)
        any = any || p;
    if (any) {
        // attached penalties on the second stream, beside the image cost:
        // forked after whatever wrote the parameters, joined before the
        // finalize and reduce that add them to the cost and gradient
        HIP_TRY(hipEventRecord(c->pen_fork, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->pen_stream, c->pen_fork, 0));
        for (int i = 0; i < (int)c->penalty.size(); ++i)
            if (c->penalty[i]) {
                const InstDesc &d = c->inst[i];
                const int st = mia_hip_rt::divcurl_enqueue(c->penalty[i], c->pen_stream, d_params + d.par_off,
                                                           c->penalty_weight[i], c->d_pen_grad + d.par_off,
                                                           c->d_pen_value + i);
                if (st)
                    return st;
            }
        HIP_TRY(hipEventRecord(c->pen_join, c->pen_stream));
        pl.pen = PenaltyIn{c->d_pen_on, c->d_pen_w, c->d_pen_value, c->d_pen_grad};
        pl.pen_join = c->pen_join;
    }
    return enqueue_plan(pl, c->stream, timed ? c : nullptr, d_params, d_grad, d_costs);
}

// A captured graph holds the launches of the settings it was captured with.
void drop_graph(mia_hip_ctx *c)
{
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    if (c->graph) {
        (void)hipGraphDestroy(c->graph);
        c->graph = nullptr;
    }
}

int check_ready(mia_hip_ctx *c)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    for (int i = 0; i < c->n_inst; ++i)
        if (!c->images_set[i])
            return fail(MIA_HIP_ERR_STATE, "images of instance " + std::to_string(i) + " are not set");
    return MIA_HIP_OK;
}

// Clamp range and, for cubic interpolation, the in-place prefilter of a
// moving image on stream s (the double scratch holds one causal pass: as many
// doubles as the image has voxels).
int enqueue_prefilter(const InstDesc &d, int dim, int degree, float *img, float2 *mm, float2 *clamp_out,
                      double *scratch, hipStream_t s)
{
    const long long nvox = (long long)d.nx * d.ny * d.nz;
    const int nbm = (int)std::min<long long>(1024, (nvox + kBlock * 16 - 1) / (kBlock * 16));
    hipLaunchKernelGGL(k_minmax_partial, dim3(nbm), dim3(kBlock), 0, s, img, nvox, mm);
    LAUNCH_CHECK("minmax");
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(64), 0, s, (const float2 *)mm, nbm, clamp_out);
    LAUNCH_CHECK("minmax_final");
    if (degree != 3)
        return MIA_HIP_OK;
    struct Pass {
        long long nlines;
        int n;
        long long stride;
        int mode;
    } passes[3] = {{(long long)d.ny * d.nz, d.nx, 1, 0},
                   {(long long)d.nx * d.nz, d.ny, d.nx, 1},
                   {(long long)d.nx * d.ny, d.nz, (long long)d.nx * d.ny, 2}};
    const int npass = dim == 3 ? 3 : 2;
    for (int k = 0; k < npass; ++k) {
        const Pass &p = passes[k];
        const long long blocks = (p.nlines + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_prefilter_lines, dim3(blocks), dim3(kBlock), 0, s, img, p.nlines, p.n, p.stride, p.mode,
                           d.nx, d.ny, scratch);
        LAUNCH_CHECK("prefilter");
    }
    return MIA_HIP_OK;
}

int prefilter_instance(mia_hip_ctx *c, int i)
{
    const InstDesc &d = c->inst[i];
    return enqueue_prefilter(d, c->dim, c->degree, c->d_M + d.img_off, c->d_mm, c->d_clamp + i, c->d_scratch,
                             c->stream);
}

// Stage 1 keeps one x row (plus tables and, in exact mode, a coefficient
// window) in dynamic LDS: allow more than the 64 KiB default where a batch
// whose longest row is max_nx needs it.
int set_stage1_lds(int dim, int max_nx, int max_cx, int max_ncx)
{
    for (int exact = 0; exact < 2; ++exact) {
        const size_t lds = stage1_layout(dim, max_nx, max_cx, max_ncx, exact != 0).bytes;
        if (lds > 160 * 1024)
            return fail(MIA_HIP_ERR_LIMIT, "stage-1 LDS requirement exceeds 160 KiB");
        const void *fn[2][2] = {{(const void *)k_stage1_3d<3, 0>, (const void *)k_stage1_3d<1, 0>},
                                {(const void *)k_stage1_3d<3, 1>, (const void *)k_stage1_3d<1, 1>}};
        const void *fn2[2][2] = {{(const void *)k_stage1_2d<3, 0>, (const void *)k_stage1_2d<1, 0>},
                                 {(const void *)k_stage1_2d<3, 1>, (const void *)k_stage1_2d<1, 1>}};
        const void *fnz[4] = {(const void *)k_stage1z_3d<3, kS1zRows, false>,
                              (const void *)k_stage1z_3d<1, kS1zRows, false>,
                              (const void *)k_stage1z_3d<3, kS1zRows, true>,
                              (const void *)k_stage1z_3d<1, kS1zRows, true>};
        if (lds > 64 * 1024) {
            for (int k = 0; k < 2; ++k)
                HIP_TRY(hipFuncSetAttribute(dim == 3 ? fn[exact][k] : fn2[exact][k],
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            if (dim == 3 && !exact)
                for (int k = 0; k < 4; ++k)
                    HIP_TRY(hipFuncSetAttribute(fnz[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        }
    }
    return MIA_HIP_OK;
}

int collect_profile(mia_hip_ctx *c)
{
    if (c->pending.empty())
        return MIA_HIP_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &e : c->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e.a, e.b));
        c->prof_ms[e.stage] += ms;
        c->prof_n[e.stage] += 1;
        c->pool.push_back(e.a);
        c->pool.push_back(e.b);
    }
    c->pending.clear();
    return MIA_HIP_OK;
}

} // namespace

namespace mia_hip_rt {
int instance_geometry(const mia_hip_ctx *c, int inst, int *dim, uint32_t size[3], int *device)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= c->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    const InstDesc &d = c->inst[inst];
    *dim = c->dim;
    size[0] = (uint32_t)d.nx;
    size[1] = (uint32_t)d.ny;
    size[2] = (uint32_t)d.nz;
    *device = c->device;
    return MIA_HIP_OK;
}
} // namespace mia_hip_rt

extern "C" {

const char *mia_hip_last_error(void)
{
    return mia_hip_rt::last_error();
}

int mia_hip_spline_csize(uint32_t n, float rate, int dim)
{
    if (rate < 1.0f || n < 1)
        return -1;
    float v;
    if (dim == 3)
        v = ((float)(n - 1) + rate) / rate;
    else
        v = (((float)n + rate) - 1.0f) / rate;
    const int cs = (int)v;
    return cs <= 1 ? -1 : cs + 2;
}

int mia_hip_ctx_create(int device, mia_hip_ctx **out)
{
    if (!out)
        return fail(MIA_HIP_ERR_ARG, "null output pointer");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(MIA_HIP_ERR_ARG, "no HIP device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(MIA_HIP_ERR_DEVICE, std::string("device is ") + prop.gcnArchName + ", this library targets gfx950");
    // the joint optimizer's streams first (see joint_streams)
    hipError_t e = mia_hip_rt::joint_streams(device, nullptr);
    if (e != hipSuccess)
        return fail(MIA_HIP_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    mia_hip_ctx *c = new mia_hip_ctx();
    c->device = device;
    e = mia_hip_rt::stream_acquire(device, &c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(MIA_HIP_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return MIA_HIP_OK;
}

int mia_hip_ctx_destroy(mia_hip_ctx *c)
{
    if (!c)
        return MIA_HIP_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_buffers(c);
    pool_trim(c);
    for (auto &e : c->pending) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : c->pool)
        (void)hipEventDestroy(e);
    if (c->pen_stream) {
        mia_hip_rt::stream_release(c->device, c->pen_stream);
        (void)hipEventDestroy(c->pen_fork);
        (void)hipEventDestroy(c->pen_join);
    }
    for (void *h : {(void *)c->h_params, (void *)c->h_grad, (void *)c->h_costs, (void *)c->h_done, (void *)c->h_up})
        if (h)
            (void)hipHostFree(h);
    mia_hip_rt::stream_release(c->device, c->stream);
    delete c;
    return MIA_HIP_OK;
}

namespace {
int setup_impl(mia_hip_ctx *c, int dim, int n_inst, const uint32_t *sizes, const uint32_t *csizes, int degree,
               bool keep_images);
}

int mia_hip_setup(mia_hip_ctx *c, int dim, int n_inst, const uint32_t *sizes, const uint32_t *csizes, int degree)
{
    return setup_impl(c, dim, n_inst, sizes, csizes, degree, false);
}

int mia_hip_setup_grid(mia_hip_ctx *c, const uint32_t *csizes)
{
    if (!c || !csizes)
        return fail(MIA_HIP_ERR_ARG, "null argument");
    if (c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    std::vector<uint32_t> sizes(3 * c->n_inst);
    for (int i = 0; i < c->n_inst; ++i) {
        sizes[3 * i] = (uint32_t)c->inst[i].nx;
        sizes[3 * i + 1] = (uint32_t)c->inst[i].ny;
        sizes[3 * i + 2] = (uint32_t)c->inst[i].nz;
    }
    return setup_impl(c, c->dim, c->n_inst, sizes.data(), csizes, c->degree, true);
}

namespace {
int setup_impl(mia_hip_ctx *c, int dim, int n_inst, const uint32_t *sizes, const uint32_t *csizes, int degree,
               bool keep_images)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (dim != 2 && dim != 3)
        return fail(MIA_HIP_ERR_ARG, "dim must be 2 or 3");
    if (n_inst < 1 || !sizes || !csizes)
        return fail(MIA_HIP_ERR_ARG, "need at least one instance with sizes and coefficient sizes");
    if (degree != 1 && degree != 3)
        return fail(MIA_HIP_ERR_ARG, "image interpolation kernel must be bspline:d=1 or bspline:d=3");
    if (c->force_variant >= (dim == 3 ? kW3Variants : kW2Variants))
        return fail(MIA_HIP_ERR_ARG, "forced warp variant " + std::to_string(c->force_variant) + " does not exist in " +
                                         std::to_string(dim) + "D");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // a new coefficient grid for the same images keeps the uploaded moving
    // (prefiltered) and fixed images and their clamp ranges: the image
    // buffers' layout depends on the image sizes only
    float *kept_M = nullptr, *kept_R = nullptr;
    float2 *kept_clamp = nullptr;
    if (keep_images) {
        kept_M = c->d_M;
        kept_R = c->d_R;
        kept_clamp = c->d_clamp;
        c->d_M = c->d_R = nullptr;
        c->d_clamp = nullptr;
    }
    struct KeptRelease { // released like the others if the set-up fails
        mia_hip_ctx *c;
        float *&M, *&R;
        float2 *&cl;
        ~KeptRelease()
        {
            pool_release(c, M);
            pool_release(c, R);
            pool_release(c, cl);
        }
    } kept_guard{c, kept_M, kept_R, kept_clamp};
    free_buffers(c);
    c->dim = dim;
    c->degree = degree;
    c->n_inst = 0;
    c->inst.assign(n_inst, InstDesc());
    c->s1_prefix.assign(n_inst + 1, 0);
    c->red_prefix.assign(n_inst + 1, 0);
    c->s1z_prefix.assign(n_inst + 1, 0);
    c->redz_prefix.assign(n_inst + 1, 0);
    if (!keep_images)
        c->images_set.assign(n_inst, 0);
    c->penalty.assign(n_inst, nullptr);
    c->penalty_weight.assign(n_inst, 0.0);
    c->inst_variant.assign(n_inst, 0);
    // 2D: one warp variant for the batch where the tuning table times them
    int batch_v2 = -1;
    if (dim == 2 && c->force_variant < 0 && n_inst > 1) {
        std::vector<long long> npix(n_inst);
        for (int i = 0; i < n_inst; ++i)
            npix[i] = (long long)sizes[3 * i] * sizes[3 * i + 1];
        batch_v2 = batch_warp2d_choice(c->device, degree, npix);
    }
    std::vector<float> tw;
    std::vector<int> ts, cells;
    long long img = 0, par = 0, part = 0, maxv = 0;
    int max_nx = 0, max_cx = 0, max_ncx = 0, n_zcell = 0, w2_win = 0;
    for (int i = 0; i < n_inst; ++i) {
        InstDesc &d = c->inst[i];
        d.nx = (int)sizes[3 * i];
        d.ny = (int)sizes[3 * i + 1];
        d.nz = dim == 3 ? (int)sizes[3 * i + 2] : 1;
        d.cx = (int)csizes[3 * i];
        d.cy = (int)csizes[3 * i + 1];
        d.cz = dim == 3 ? (int)csizes[3 * i + 2] : 1;
        if (d.nx > kMaxRow)
            return fail(MIA_HIP_ERR_LIMIT, "image x size " + std::to_string(d.nx) + " exceeds the limit " +
                                               std::to_string(kMaxRow));
        // the 3D box fill and z-cell stage 1 address an image with 32-bit byte offsets
        if (dim == 3 && (long long)d.nx * d.ny * d.nz * (long long)sizeof(float) >= (1LL << 31))
            return fail(MIA_HIP_ERR_LIMIT, "3D image of instance " + std::to_string(i) +
                                               " has 2^29 voxels or more (limit of the 32-bit image offsets)");
        AxisTable ax, ay, az;
        int st = make_axis_table(d.nx, d.cx, ax, "x");
        if (st)
            return st;
        st = make_axis_table(d.ny, d.cy, ay, "y");
        if (st)
            return st;
        if (dim == 3) {
            st = make_axis_table(d.nz, d.cz, az, "z");
            if (st)
                return st;
        } else {
            az.w = {1.0f, 0.0f, 0.0f, 0.0f};
            az.start = {0};
            az.cells = {0, 1};
            az.nbands = 1;
        }
        d.ncx = ax.nbands;
        d.ncy = ay.nbands;
        d.ncz = az.nbands;
        int band_rows = 0;
        for (int k = 0; k < d.ncy; ++k)
            band_rows = std::max(band_rows, ay.cells[k + 1] - ay.cells[k]);
        // reference-gradient stage 1 per (z knot cell, y band) where the
        // instance allows it and the tuned layout (or the forced one) asks
        // for it: decided per instance from its sizes, so that an instance
        // computes the same in any batch
        const long long nvox = (long long)d.nx * d.ny * d.nz;
        const long long ncoef = (long long)d.cx * d.cy * d.cz;
        d.zcell = zcell_eligible(dim, d.nx, band_rows, ncoef) &&
                          (c->force_s1 >= 0 ? c->force_s1 : tuned_choice(c->device, MIA_HIP_TUNE_STAGE1, degree, nvox)) == 1
                      ? 1
                      : 0;
        n_zcell += d.zcell;
        AxisTable *axes[3] = {&ax, &ay, &az};
        int *toff[3] = {&d.tab_x, &d.tab_y, &d.tab_z};
        int *coff[3] = {&d.cell_x, &d.cell_y, &d.cell_z};
        for (int a = 0; a < 3; ++a) {
            *toff[a] = (int)ts.size();
            ts.insert(ts.end(), axes[a]->start.begin(), axes[a]->start.end());
            tw.insert(tw.end(), axes[a]->w.begin(), axes[a]->w.end());
            *coff[a] = (int)cells.size();
            cells.insert(cells.end(), axes[a]->cells.begin(), axes[a]->cells.end());
        }
        d.img_off = img;
        img += align64(nvox);
        d.par_off = par;
        par += dim * ncoef;
        d.part_off = part;
        // stage-1 partial sums: per (slice, y band) row of 12 cx (exact
        // gradient) or per (z knot cell, y band) row of 48 cx (reference)
        part += dim == 3 ? std::max((long long)d.nz * d.ncy * 12 * d.cx, (long long)d.ncz * d.ncy * 48 * d.cx)
                         : (long long)d.ncy * 8 * d.cx;
        maxv = std::max(maxv, nvox);
        max_nx = std::max(max_nx, d.nx);
        if (dim == 2) {
            // the 2D warp tile and the coefficient window it loads
            const int v = c->force_variant >= 0 ? c->force_variant
                          : batch_v2 >= 0      ? batch_v2
                                               : tuned_choice(c->device, MIA_HIP_TUNE_WARP2D, degree, nvox);
            const int wx = tile_span(d.nx, w2_tile_x(v), [&](int p) { return ax.start[p]; });
            const int wy = tile_span(d.ny, w2_tile_y(v), [&](int p) { return ay.start[p]; });
            w2_win = std::max(w2_win, wx * wy);
            c->inst_variant[i] = v;
        }
        max_cx = std::max(max_cx, d.cx);
        max_ncx = std::max(max_ncx, d.ncx);
        c->s1_prefix[i + 1] = c->s1_prefix[i] + (dim == 3 ? d.nz * d.ncy : d.ncy);
        // 3D reduce: 64 gradient components per block (4 waves split the z
        // support); 2D: 256 per block
        const int per_block = dim == 3 ? 64 : kBlock;
        c->red_prefix[i + 1] = c->red_prefix[i] + (int)((dim * ncoef + per_block - 1) / per_block);
        c->s1z_prefix[i + 1] = c->s1z_prefix[i] + d.ncz * d.ncy;
        c->redz_prefix[i + 1] = c->redz_prefix[i] + (int)((dim * ncoef + kBlock - 1) / kBlock);
    }
    // warp groups: instances by tile variant (tuned per size bucket unless
    // forced)
    c->inst_group.assign(n_inst, 0);
    c->inst_slot.assign(n_inst, 0);
    c->wgroups.clear();
    {
        int gidx[kMaxTuneChoices];
        for (int v = 0; v < kMaxTuneChoices; ++v)
            gidx[v] = -1;
        for (int i = 0; i < n_inst; ++i) {
            const InstDesc &d = c->inst[i];
            const long long nvox = (long long)d.nx * d.ny * d.nz;
            const int v = dim == 3 ? (c->force_variant >= 0 ? c->force_variant : tuned_choice(c->device, MIA_HIP_TUNE_WARP3D, degree, nvox))
                                   : c->inst_variant[i];
            c->inst_variant[i] = v;
            if (gidx[v] < 0) {
                gidx[v] = (int)c->wgroups.size();
                c->wgroups.emplace_back();
                c->wgroups.back().variant = v;
                c->wgroups.back().prefix.push_back(0);
            }
            auto &G = c->wgroups[gidx[v]];
            int tiles;
            if (dim == 3) {
                int tx, ty, tz;
                w3_tile_dims(v, tx, ty, tz);
                tiles = ((d.nx + tx - 1) / tx) * ((d.ny + ty - 1) / ty) * ((d.nz + tz - 1) / tz);
            } else {
                const int tx = w2_tile_x(v), ty = w2_tile_y(v);
                tiles = ((d.nx + tx - 1) / tx) * ((d.ny + ty - 1) / ty);
            }
            c->inst_group[i] = gidx[v];
            c->inst_slot[i] = (int)G.ids.size();
            G.ids.push_back(i);
            G.prefix.push_back(G.prefix.back() + tiles);
        }
    }
    std::vector<int> wg;
    for (auto &G : c->wgroups) {
        G.ids_off = (int)wg.size();
        wg.insert(wg.end(), G.ids.begin(), G.ids.end());
        G.prefix_off = (int)wg.size();
        wg.insert(wg.end(), G.prefix.begin(), G.prefix.end());
    }
    c->s1tab_off = -1;
    c->redtab_off = -1;
    if (dim == 2) {
        // reduce: each gradient block's {instance, block of the instance}
        while (wg.size() % 2)
            wg.push_back(0);
        c->redtab_off = (int)wg.size();
        for (int i = 0; i < n_inst; ++i)
            for (int k = 0; k < c->red_prefix[i + 1] - c->red_prefix[i]; ++k) {
                wg.push_back(i);
                wg.push_back(k);
            }
        // stage 1: each block's {instance, y band, first row, end row}
        while (wg.size() % 4)
            wg.push_back(0);
        c->s1tab_off = (int)wg.size();
        for (int i = 0; i < n_inst; ++i) {
            const InstDesc &d = c->inst[i];
            for (int band = 0; band < c->s1_prefix[i + 1] - c->s1_prefix[i]; ++band) {
                const int e[4] = {i, band, cells[d.cell_y + band], cells[d.cell_y + band + 1]};
                wg.insert(wg.end(), e, e + 4);
            }
        }
    }
    if (dim == 2)
        for (auto &G : c->wgroups) {
            // each tile's instance, origin and coefficient window, as the
            // kernel would derive them from the tile prefix and the axis tables
            while (wg.size() % 4)
                wg.push_back(0);
            G.tiles2d_off = (int)wg.size();
            const int tx = w2_tile_x(G.variant), ty = w2_tile_y(G.variant);
            for (int i : G.ids) {
                const InstDesc &d = c->inst[i];
                const int tiles_x = (d.nx + tx - 1) / tx, tiles_y = (d.ny + ty - 1) / ty;
                for (int lb = 0; lb < tiles_x * tiles_y; ++lb) {
                    const int x0 = (lb % tiles_x) * tx, y0 = (lb / tiles_x) * ty;
                    const int kx0 = ts[d.tab_x + x0], ky0 = ts[d.tab_y + y0];
                    const int wxn = ts[d.tab_x + std::min(x0 + tx, d.nx) - 1] + 4 - kx0;
                    const int wyn = ts[d.tab_y + std::min(y0 + ty, d.ny) - 1] + 4 - ky0;
                    const int e[8] = {i, x0, y0, kx0, ky0, wxn, wyn, 0};
                    wg.insert(wg.end(), e, e + 8);
                }
            }
        }
    c->n_img = img;
    c->n_par = par;
    c->n_part = part;
    c->max_vox = maxv;
    c->max_nx = max_nx;
    if (w2_win > kW2Win)
        return fail(MIA_HIP_ERR_ARG, "2D coefficient window exceeds its bound (spline rate below 1?)");
    c->w2_win = w2_win;
    c->max_cx = max_cx;
    c->max_ncx = max_ncx;
    c->n_zcell = n_zcell;
    {
        const int st = set_stage1_lds(dim, max_nx, max_cx, max_ncx);
        if (st)
            return st;
    }
    if (keep_images) {
        c->d_M = kept_M;
        c->d_R = kept_R;
        c->d_clamp = kept_clamp;
        kept_M = kept_R = nullptr; // owned by the context again
        kept_clamp = nullptr;
    }
    int st;
    if ((st = dev_alloc(c, &c->d_inst, n_inst)) || (st = dev_alloc(c, &c->d_wgroups, wg.size())) ||
        (st = dev_alloc(c, &c->d_s1_prefix, n_inst + 1)) || (st = dev_alloc(c, &c->d_red_prefix, n_inst + 1)) ||
        (st = dev_alloc(c, &c->d_s1z_prefix, n_inst + 1)) || (st = dev_alloc(c, &c->d_redz_prefix, n_inst + 1)) ||
        (st = keep_images ? 0 : dev_alloc(c, &c->d_M, img)) || (st = keep_images ? 0 : dev_alloc(c, &c->d_R, img)) ||
        (st = dev_alloc(c, &c->d_W, img)) || (st = dev_alloc(c, &c->d_P, part)) ||
        (st = keep_images ? 0 : dev_alloc(c, &c->d_clamp, n_inst)) || (st = dev_alloc(c, &c->d_mm, 1024)) ||
        (st = dev_alloc(c, &c->d_tw, tw.size() / 4)) ||
        (st = dev_alloc(c, &c->d_ts, ts.size())) || (st = dev_alloc(c, &c->d_cells, cells.size())) ||
        (st = dev_alloc(c, &c->d_ssd, c->s1_prefix[n_inst] + c->s1z_prefix[n_inst])) ||
        (st = dev_alloc(c, &c->d_cnt, c->s1_prefix[n_inst] + c->s1z_prefix[n_inst])) ||
        (st = dev_alloc(c, &c->d_gscale, n_inst)) || (st = dev_alloc(c, &c->d_params, par)) ||
        (st = dev_alloc(c, &c->d_grad, par)) || (st = dev_alloc(c, &c->d_costs, n_inst)) ||
        (st = dev_alloc(c, &c->d_scratch, maxv))) {
        free_buffers(c);
        return st;
    }
    if ((st = host_reserve(&c->h_params, &c->cap_params, (size_t)par)) ||
        (st = host_reserve(&c->h_grad, &c->cap_grad, (size_t)par)) ||
        (st = host_reserve(&c->h_costs, &c->cap_costs, (size_t)n_inst)))
        return st;
    if ((st = stage_upload(c, c->d_inst, c->inst.data(), sizeof(InstDesc) * n_inst)) ||
        (st = stage_upload(c, c->d_wgroups, wg.data(), sizeof(int) * wg.size())) ||
        (st = stage_upload(c, c->d_s1_prefix, c->s1_prefix.data(), sizeof(int) * (n_inst + 1))) ||
        (st = stage_upload(c, c->d_red_prefix, c->red_prefix.data(), sizeof(int) * (n_inst + 1))) ||
        (st = stage_upload(c, c->d_s1z_prefix, c->s1z_prefix.data(), sizeof(int) * (n_inst + 1))) ||
        (st = stage_upload(c, c->d_redz_prefix, c->redz_prefix.data(), sizeof(int) * (n_inst + 1))) ||
        (st = stage_upload(c, c->d_tw, tw.data(), sizeof(float) * tw.size())) ||
        (st = stage_upload(c, c->d_ts, ts.data(), sizeof(int) * ts.size())) ||
        (st = stage_upload(c, c->d_cells, cells.data(), sizeof(int) * cells.size())))
        return st;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->n_inst = n_inst;
    return MIA_HIP_OK;
}
} // namespace

static int set_images_impl(mia_hip_ctx *c, int inst, const void *moving, const void *fixed, hipMemcpyKind kind)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= c->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    if (!moving || !fixed)
        return fail(MIA_HIP_ERR_ARG, "null image pointer");
    HIP_TRY(hipSetDevice(c->device));
    const InstDesc &d = c->inst[inst];
    const size_t bytes = sizeof(float) * (size_t)d.nx * d.ny * d.nz;
    HIP_TRY(hipMemcpyAsync(c->d_M + d.img_off, moving, bytes, kind, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_R + d.img_off, fixed, bytes, kind, c->stream));
    int st = prefilter_instance(c, inst);
    if (st)
        return st;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->images_set[inst] = 1;
    return MIA_HIP_OK;
}

int mia_hip_set_images(mia_hip_ctx *c, int inst, const float *moving, const float *fixed)
{
    return set_images_impl(c, inst, moving, fixed, hipMemcpyHostToDevice);
}

int mia_hip_set_images_device(mia_hip_ctx *c, int inst, const void *d_moving, const void *d_fixed)
{
    return set_images_impl(c, inst, d_moving, d_fixed, hipMemcpyDeviceToDevice);
}

int mia_hip_set_cost(mia_hip_ctx *c, int norm, float autothresh, double weight, int grad_mode)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (!(autothresh >= 0.0f && autothresh <= 1000.0f))
        return fail(MIA_HIP_ERR_ARG, "autothresh must be in [0, 1000]");
    if (grad_mode != MIA_HIP_GRAD_REFERENCE && grad_mode != MIA_HIP_GRAD_EXACT)
        return fail(MIA_HIP_ERR_ARG, "unknown gradient mode");
    c->norm = norm ? 1 : 0;
    c->thr = autothresh;
    c->weight = weight;
    c->grad_mode = grad_mode;
    drop_graph(c);
    return MIA_HIP_OK;
}

int mia_hip_n_params(mia_hip_ctx *c, int inst, size_t *n, size_t *offset)
{
    if (!c || c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= c->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    const InstDesc &d = c->inst[inst];
    if (n)
        *n = (size_t)c->dim * d.cx * d.cy * d.cz;
    if (offset)
        *offset = (size_t)d.par_off;
    return MIA_HIP_OK;
}

int mia_hip_set_penalty(mia_hip_ctx *c, int inst, mia_hip_divcurl *p, double weight)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= c->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    if (p) {
        if (c->dim != 3)
            return fail(MIA_HIP_ERR_ARG, "set_penalty: the divcurl penalty is 3D");
        uint32_t cs[3];
        int st = mia_hip_rt::divcurl_grid(p, cs);
        if (st)
            return st;
        const InstDesc &d = c->inst[inst];
        if ((int)cs[0] != d.cx || (int)cs[1] != d.cy || (int)cs[2] != d.cz)
            return fail(MIA_HIP_ERR_SIZE, "set_penalty: the penalty's coefficient grid differs from the instance's");
        if (!std::isfinite(weight))
            return fail(MIA_HIP_ERR_ARG, "set_penalty: weight must be finite");
    }
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (p && !c->d_pen_on) {
        int st;
        if ((st = dev_alloc(c, &c->d_pen_on, c->n_inst)) || (st = dev_alloc(c, &c->d_pen_w, c->n_inst)) ||
            (st = dev_alloc(c, &c->d_pen_value, c->n_inst)) || (st = dev_alloc(c, &c->d_pen_grad, c->n_par)))
            return st;
    }
    if (p && !c->pen_stream) {
        HIP_TRY(mia_hip_rt::stream_acquire(c->device, &c->pen_stream));
        HIP_TRY(hipEventCreateWithFlags(&c->pen_fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->pen_join, hipEventDisableTiming));
    }
    c->penalty[inst] = p;
    c->penalty_weight[inst] = p ? weight : 0.0;
    if (c->d_pen_on) {
        std::vector<int> on(c->n_inst);
        for (int i = 0; i < c->n_inst; ++i)
            on[i] = c->penalty[i] ? 1 : 0;
        int st;
        if ((st = stage_upload(c, c->d_pen_on, on.data(), sizeof(int) * c->n_inst)) ||
            (st = stage_upload(c, c->d_pen_w, c->penalty_weight.data(), sizeof(double) * c->n_inst)))
            return st;
    }
    drop_graph(c);
    return MIA_HIP_OK;
}


int mia_hip_get_stage1_path(mia_hip_ctx *c, int inst, int *path)
{
    if (!c || !path)
        return fail(MIA_HIP_ERR_ARG, "null argument");
    if (c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= c->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    // the exact gradient always runs per slice
    *path = c->inst[inst].zcell && c->grad_mode != MIA_HIP_GRAD_EXACT ? 1 : 0;
    return MIA_HIP_OK;
}

int mia_hip_cost_grad(mia_hip_ctx *c, const double *params, double *grad, double *costs)
{
    int st = check_ready(c);
    if (st)
        return st;
    if (!params || !grad || !costs)
        return fail(MIA_HIP_ERR_ARG, "null pointer");
    HIP_TRY(hipSetDevice(c->device));
    std::memcpy(c->h_params, params, sizeof(double) * c->n_par);
    HIP_TRY(hipMemcpyAsync(c->d_params, c->h_params, sizeof(double) * c->n_par, hipMemcpyHostToDevice, c->stream));
    if ((st = enqueue_eval(c, c->d_params, c->d_grad, c->d_costs, c->profile)))
        return st;
    HIP_TRY(hipMemcpyAsync(c->h_grad, c->d_grad, sizeof(double) * c->n_par, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_costs, c->d_costs, sizeof(double) * c->n_inst, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(grad, c->h_grad, sizeof(double) * c->n_par);
    std::memcpy(costs, c->h_costs, sizeof(double) * c->n_inst);
    return MIA_HIP_OK;
}

int mia_hip_cost_grad_device(mia_hip_ctx *c, const void *d_params, void *d_grad, void *d_costs)
{
    int st = check_ready(c);
    if (st)
        return st;
    if (!d_params || !d_grad || !d_costs)
        return fail(MIA_HIP_ERR_ARG, "null pointer");
    HIP_TRY(hipSetDevice(c->device));
    return enqueue_eval(c, (const double *)d_params, (double *)d_grad, (double *)d_costs, c->profile);
}

int mia_hip_device_buffers(mia_hip_ctx *c, void **d_params, void **d_grad, void **d_costs)
{
    if (!c || c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (d_params)
        *d_params = c->d_params;
    if (d_grad)
        *d_grad = c->d_grad;
    if (d_costs)
        *d_costs = c->d_costs;
    return MIA_HIP_OK;
}

int mia_hip_warp(mia_hip_ctx *c, int inst, const double *params_inst, float *out)
{
    if (!c || c->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= c->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    if (!c->images_set[inst])
        return fail(MIA_HIP_ERR_STATE, "images of instance " + std::to_string(inst) + " are not set");
    if (!params_inst || !out)
        return fail(MIA_HIP_ERR_ARG, "null pointer");
    HIP_TRY(hipSetDevice(c->device));
    const InstDesc &d = c->inst[inst];
    const long long ncoef = (long long)d.cx * d.cy * d.cz;
    std::memcpy(c->h_params + d.par_off, params_inst, sizeof(double) * c->dim * ncoef);
    HIP_TRY(hipMemcpyAsync(c->d_params + d.par_off, c->h_params + d.par_off, sizeof(double) * c->dim * ncoef,
                           hipMemcpyHostToDevice, c->stream));
    const auto &G = c->wgroups[c->inst_group[inst]];
    const int boff = G.prefix[c->inst_slot[inst]];
    int st = launch_warp(plan_of(c), c->stream, nullptr, c->d_params, c->inst_group[inst], boff,
                         G.prefix[c->inst_slot[inst] + 1] - boff);
    if (st)
        return st;
    HIP_TRY(hipMemcpyAsync(out, c->d_W + d.img_off, sizeof(float) * (size_t)d.nx * d.ny * d.nz, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MIA_HIP_OK;
}

int mia_hip_graph_capture(mia_hip_ctx *c)
{
    int st = check_ready(c);
    if (st)
        return st;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    if (c->graph) {
        (void)hipGraphDestroy(c->graph);
        c->graph = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(c->d_params, c->h_params, sizeof(double) * c->n_par, hipMemcpyHostToDevice,
                                  c->stream);
    int rc = e == hipSuccess ? enqueue_eval(c, c->d_params, c->d_grad, c->d_costs, false)
                             : fail(MIA_HIP_ERR_DEVICE, hipGetErrorString(e));
    if (rc == MIA_HIP_OK) {
        e = hipMemcpyAsync(c->h_grad, c->d_grad, sizeof(double) * c->n_par, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->h_costs, c->d_costs, sizeof(double) * c->n_inst, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess)
            rc = fail(MIA_HIP_ERR_DEVICE, hipGetErrorString(e));
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc != MIA_HIP_OK) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    if (e2 != hipSuccess)
        return fail(MIA_HIP_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
    c->graph = g;
    HIP_TRY(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    return MIA_HIP_OK;
}

int mia_hip_graph_params(mia_hip_ctx *c, double **params)
{
    if (!c || !c->gexec)
        return fail(MIA_HIP_ERR_STATE, "no captured graph");
    *params = c->h_params;
    return MIA_HIP_OK;
}

int mia_hip_graph_outputs(mia_hip_ctx *c, double **grad, double **costs)
{
    if (!c || !c->gexec)
        return fail(MIA_HIP_ERR_STATE, "no captured graph");
    *grad = c->h_grad;
    *costs = c->h_costs;
    return MIA_HIP_OK;
}

int mia_hip_graph_run(mia_hip_ctx *c)
{
    if (!c || !c->gexec)
        return fail(MIA_HIP_ERR_STATE, "no captured graph");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MIA_HIP_OK;
}

namespace {

// One context's device L-BFGS: the step kernels it needs, its chunk of
// iterations and its results (mia_hip_optimize, mia_hip_optimize_many).
struct OptRun {
    mia_hip_ctx *c = nullptr;
    int maxeval = 0, memory = 0, ipg = 1;
    double first_step = 1.0;
    std::vector<int> seg; // segment prefix of the large instances' passes
    int n_small = 0;      // instances stepped by one block each
    bool paired = true;   // every instance's vectors 16-byte aligned, even lengths
    OptBuffers ob{};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool active = true;
};

// Optimizer buffers (re)allocated for `memory`, the starting point and the
// initial state uploaded on the context's stream.
int opt_run_prepare(OptRun &r, mia_hip_ctx *c, const double *x0, int maxeval, int memory, double first_step,
                    int iters_per_graph)
{
    int st = check_ready(c);
    if (st)
        return st;
    if (!x0)
        return fail(MIA_HIP_ERR_ARG, "null starting point");
    if (maxeval < 1 || memory < 1 || memory > kMaxMemory || !(first_step > 0.0) || iters_per_graph < 1)
        return fail(MIA_HIP_ERR_ARG, "optimize: need maxeval >= 1, 1 <= memory <= 16, first_step > 0, "
                                     "iters_per_graph >= 1");
    r.c = c;
    r.maxeval = maxeval;
    r.memory = memory;
    r.first_step = first_step;
    r.ipg = iters_per_graph;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int n = c->n_inst;
    // instances of up to kLbRegN parameters step in one block each
    // (k_lbfgs_step_small), larger ones in segments of kSeg parameters
    // (k_lbfgs_pass1..3): the segment prefix counts none for the small ones
    r.seg.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const InstDesc &d = c->inst[i];
        const long long np = (long long)c->dim * d.cx * d.cy * d.cz;
        const bool small = np <= kLbRegN;
        r.n_small += small;
        r.paired = r.paired && np % 2 == 0 && d.par_off % 2 == 0;
        r.seg[i + 1] = r.seg[i] + (small ? 0 : (int)((np + kSeg - 1) / kSeg));
    }
    const std::vector<int> &seg = r.seg;
    if (c->opt_memory != memory) {
        void *old[] = {c->d_oX, c->d_oG, c->d_oD, c->d_oBX, c->d_oS, c->d_oY, c->d_obest,
                       c->d_ost, c->d_oSY, c->d_oYY, c->d_oPART, c->d_oPART2, c->d_oseg, c->d_ost2};
        for (void *p : old)
            pool_release(c, p);
        c->opt_memory = 0;
        const size_t R = (size_t)memory + 1; // ring: memory pairs and the candidate's slot
        if ((st = dev_alloc(c, &c->d_oX, c->n_par)) || (st = dev_alloc(c, &c->d_oG, c->n_par)) ||
            (st = dev_alloc(c, &c->d_oD, c->n_par)) || (st = dev_alloc(c, &c->d_oBX, c->n_par)) ||
            (st = dev_alloc(c, &c->d_oS, R * c->n_par)) || (st = dev_alloc(c, &c->d_oY, R * c->n_par)) ||
            (st = dev_alloc(c, &c->d_obest, n)) || (st = dev_alloc(c, &c->d_ost, n)) ||
            (st = dev_alloc(c, &c->d_oSY, R * R * n)) || (st = dev_alloc(c, &c->d_oYY, R * R * n)) ||
            (st = dev_alloc(c, &c->d_oPART, (size_t)kPart * seg[n])) ||
            (st = dev_alloc(c, &c->d_oPART2, (size_t)kPart * seg[n])) || (st = dev_alloc(c, &c->d_ost2, n)) ||
            (st = dev_alloc(c, &c->d_oseg, n + 1)))
            return st;
        c->opt_memory = memory;
        c->opt_segments = seg[n];
    }
    // initial state: parameters, phase 0, not done (plain copies, outside the graph)
    std::vector<OptState> st0(n);
    for (auto &o : st0) {
        o = OptState();
        o.phase = 0;
        o.mode = 0;
    }
    if ((st = stage_upload(c, c->d_params, x0, sizeof(double) * c->n_par)) ||
        (st = stage_upload(c, c->d_ost, st0.data(), sizeof(OptState) * n)) ||
        (st = stage_upload(c, c->d_oseg, seg.data(), sizeof(int) * (n + 1))))
        return st;
    // the done flags in pinned host memory, cleared here (the stream is idle:
    // no kernel of an earlier call still writes them), written by the kernels
    // through their device address
    if ((st = host_reserve(&c->h_done, &c->cap_done, (size_t)n)))
        return st;
    std::memset(c->h_done, 0, sizeof(int) * n);
    int *done_dev = nullptr;
    HIP_TRY(hipHostGetDevicePointer((void **)&done_dev, c->h_done, 0));
    r.ob = OptBuffers{c->d_oX,  c->d_oG,   c->d_oD,   c->d_oBX,   c->d_oS,    c->d_oY,     c->d_obest, done_dev,
                      c->d_ost, c->d_ost2, c->d_oSY,  c->d_oYY,   c->d_oPART, c->d_oPART2, c->d_oseg};
    return MIA_HIP_OK;
}

// One chunk: ipg iterations (evaluation, then the step's kernels).
int opt_run_enqueue_chunk(OptRun &r)
{
    mia_hip_ctx *c = r.c;
    const int n = c->n_inst;
    const std::vector<int> &seg = r.seg;
    const OptBuffers &ob = r.ob;
    int rc = MIA_HIP_OK;
    for (int it = 0; it < r.ipg && rc == MIA_HIP_OK; ++it) {
        rc = enqueue_eval(c, c->d_params, c->d_grad, c->d_costs, false);
        if (rc != MIA_HIP_OK)
            break;
        hipError_t e = hipSuccess;
        if (r.n_small > 0) {
            // memory <= kLbRegMem: the stored pairs stay in registers for the whole step;
            // paired: every instance's vectors start 16-byte aligned and have even
            // lengths, so each thread moves adjacent pairs of elements
            const bool reg = r.memory <= kLbRegMem;
            auto kstep = reg ? (r.paired ? k_lbfgs_step_small<true, true> : k_lbfgs_step_small<true, false>)
                             : (r.paired ? k_lbfgs_step_small<false, true> : k_lbfgs_step_small<false, false>);
            hipLaunchKernelGGL(kstep, dim3(n), dim3(kOptBlock), 0, c->stream, (const InstDesc *)c->d_inst, c->dim,
                               ob, c->d_params, (const double *)c->d_grad, (const double *)c->d_costs, r.memory,
                               r.maxeval, r.first_step);
            e = hipGetLastError();
        }
        if (e == hipSuccess && seg[n] > 0) {
            hipLaunchKernelGGL(k_lbfgs_pass1, dim3(seg[n]), dim3(kSegT), 0, c->stream, (const InstDesc *)c->d_inst, n,
                               c->dim, ob, c->d_params, (const double *)c->d_grad, (const double *)c->d_costs,
                               r.memory, r.maxeval);
            e = hipGetLastError();
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_lbfgs_pass2, dim3(seg[n]), dim3(kSegT), 0, c->stream,
                                   (const InstDesc *)c->d_inst, n, c->dim, ob, c->d_params,
                                   (const double *)c->d_costs, r.memory, r.maxeval, r.first_step);
                e = hipGetLastError();
            }
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_lbfgs_pass3, dim3(seg[n]), dim3(kSegT), 0, c->stream,
                                   (const InstDesc *)c->d_inst, n, c->dim, ob, c->d_params, r.first_step);
                e = hipGetLastError();
            }
        }
        if (e != hipSuccess)
            rc = fail(MIA_HIP_ERR_DEVICE, std::string("launch lbfgs passes: ") + hipGetErrorString(e));
    }
    return rc;
}

// Every instance done, as far as the kernels have written their flags (read
// after a chunk's event: at least that chunk's flags; a kernel of the chunk
// queued after it may already have set more)
bool opt_run_all_done(const OptRun &r)
{
    const volatile int *done = r.c->h_done;
    for (int i = 0; i < r.c->n_inst; ++i)
        if (!done[i])
            return false;
    return true;
}

// every iteration performs one evaluation per unfinished instance, and an
// instance finishes at the latest with its maxeval-th evaluation
int opt_max_launches(int maxeval, int ipg)
{
    return (maxeval + ipg - 1) / ipg + 1;
}

// Stream launches of a run on its context's stream: one chunk queued ahead
// while the host reads the previous chunk's done flags (only chunks that start
// before the maxeval-th iteration are queued ahead; a chunk queued after all of
// the run's instances finished changes nothing: finished instances' steps
// return at once).
int opt_stream_loop(OptRun &r, double *seconds)
{
    const int max_launches = opt_max_launches(r.maxeval, r.ipg);
    const int sure_launches = max_launches - 1;
    int rc = MIA_HIP_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == MIA_HIP_OK)
            rc = fail(MIA_HIP_ERR_DEVICE, std::string("optimize: ") + hipGetErrorString(e));
        return rc == MIA_HIP_OK;
    };
    for (auto &e : r.ev)
        if (!check(hipEventCreateWithFlags(&e, hipEventDisableTiming)))
            break;
    auto queue = [&](int slot) {
        if (rc == MIA_HIP_OK)
            rc = opt_run_enqueue_chunk(r);
        if (rc == MIA_HIP_OK)
            check(hipEventRecord(r.ev[slot], r.c->stream));
    };
    const auto t0 = std::chrono::steady_clock::now();
    queue(0);
    for (int l = 0; l < max_launches && rc == MIA_HIP_OK; ++l) {
        const bool ahead = l + 1 < sure_launches;
        if (ahead)
            queue((l + 1) & 1);
        if (!check(hipEventSynchronize(r.ev[l & 1])))
            break;
        if (opt_run_all_done(r) || l + 1 == max_launches)
            break;
        if (!ahead)
            queue((l + 1) & 1);
    }
    check(hipStreamSynchronize(r.c->stream));
    const auto t1 = std::chrono::steady_clock::now();
    for (auto &e : r.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (rc == MIA_HIP_OK && seconds)
        *seconds = std::chrono::duration<double>(t1 - t0).count();
    return rc;
}

// One captured graph of a chunk, launched until every instance is done.
int opt_graph_loop(OptRun &r, double *seconds)
{
    mia_hip_ctx *c = r.c;
    const int max_launches = opt_max_launches(r.maxeval, r.ipg);
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = opt_run_enqueue_chunk(r);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc != MIA_HIP_OK) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    if (e2 != hipSuccess)
        return fail(MIA_HIP_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
    hipError_t e3 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(MIA_HIP_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int l = 0; l < max_launches && rc == MIA_HIP_OK; ++l) {
        hipError_t e = hipGraphLaunch(ge, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            rc = fail(MIA_HIP_ERR_DEVICE, std::string("optimize graph: ") + hipGetErrorString(e));
            break;
        }
        if (opt_run_all_done(r))
            break;
    }
    const auto t1 = std::chrono::steady_clock::now();
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (rc == MIA_HIP_OK && seconds)
        *seconds = std::chrono::duration<double>(t1 - t0).count();
    return rc;
}

int opt_run_results(OptRun &r, double *x_out, double *costs_out, int *evals_out)
{
    mia_hip_ctx *c = r.c;
    const int n = c->n_inst;
    int st;
    if (x_out && (st = stage_download(c, x_out, c->d_params, sizeof(double) * c->n_par)))
        return st;
    if (costs_out && (st = stage_download(c, costs_out, c->d_obest, sizeof(double) * n)))
        return st;
    if (evals_out) {
        std::vector<OptState> stv(n);
        if ((st = stage_download(c, stv.data(), c->d_ost, sizeof(OptState) * n)))
            return st;
        for (int i = 0; i < n; ++i)
            evals_out[i] = stv[i].evals;
    }
    return MIA_HIP_OK;
}

// While alive, the listed contexts queue their work on other streams; on
// destruction those streams are drained and the contexts get their own back.
struct StreamSwap {
    std::vector<std::pair<mia_hip_ctx *, hipStream_t>> saved;
    ~StreamSwap()
    {
        for (auto &p : saved) {
            (void)hipStreamSynchronize(p.first->stream);
            p.first->stream = p.second;
        }
    }
};

} // namespace

int mia_hip_optimize(mia_hip_ctx *c, const double *x0, int maxeval, int memory, double first_step,
                     int iters_per_graph, double *x_out, double *costs_out, int *evals_out, double *seconds)
{
    OptRun run;
    int st = opt_run_prepare(run, c, x0, maxeval, memory, first_step, iters_per_graph);
    if (st)
        return st;
    st = c->opt_launch == MIA_HIP_LAUNCH_STREAM ? opt_stream_loop(run, seconds) : opt_graph_loop(run, seconds);
    if (st)
        return st;
    return opt_run_results(run, x_out, costs_out, evals_out);
}

int mia_hip_optimize_many(mia_hip_ctx *const *ctx, int n_ctx, const double *const *x0, int maxeval, int memory,
                          double first_step, int iters_per_graph, double *const *x_out, double *const *costs_out,
                          int *const *evals_out, double *seconds)
{
    if (!ctx || n_ctx < 1 || !x0)
        return fail(MIA_HIP_ERR_ARG, "optimize_many: need at least one context and its starting point");
    for (int k = 0; k < n_ctx; ++k) {
        if (!ctx[k])
            return fail(MIA_HIP_ERR_ARG, "optimize_many: null context " + std::to_string(k));
        if (ctx[k]->device != ctx[0]->device)
            return fail(MIA_HIP_ERR_ARG, "optimize_many: the contexts must be on one device");
        for (int j = 0; j < k; ++j)
            if (ctx[j] == ctx[k])
                return fail(MIA_HIP_ERR_ARG, "optimize_many: context " + std::to_string(k) + " is listed twice");
    }
    HIP_TRY(hipSetDevice(ctx[0]->device));
    std::array<hipStream_t, mia_hip_rt::kJointStreams> js{};
    HIP_TRY(mia_hip_rt::joint_streams(ctx[0]->device, &js));
    std::vector<OptRun> runs(n_ctx);
    int st = MIA_HIP_OK;
    {
        StreamSwap swap;
        for (int k = 0; k < n_ctx && st == MIA_HIP_OK; ++k) {
            // the context's own stream drained, then its work goes to a joint stream
            HIP_TRY(hipStreamSynchronize(ctx[k]->stream));
            swap.saved.emplace_back(ctx[k], ctx[k]->stream);
            ctx[k]->stream = js[k % mia_hip_rt::kJointStreams];
            st = opt_run_prepare(runs[k], ctx[k], x0[k], maxeval, memory, first_step, iters_per_graph);
        }
        if (st == MIA_HIP_OK) {
            // one host thread per run, each queueing its own chunks (queued
            // from one thread, the launches of several runs lag behind the
            // GPU: a 64-frame 2D series in two runs took 4.48-4.76 ms that
            // way against 4.38-4.41 ms with a thread per run)
            std::vector<int> rcs(n_ctx, MIA_HIP_OK);
            std::vector<std::string> msgs(n_ctx);
            auto work = [&](int k) {
                hipError_t e = hipSetDevice(ctx[k]->device);
                rcs[k] = e == hipSuccess ? opt_stream_loop(runs[k], nullptr)
                                         : fail(MIA_HIP_ERR_DEVICE, std::string("hipSetDevice: ") +
                                                                        hipGetErrorString(e));
                if (rcs[k] != MIA_HIP_OK)
                    msgs[k] = mia_hip_rt::last_error();
            };
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<std::thread> th;
            for (int k = 1; k < n_ctx; ++k)
                th.emplace_back(work, k);
            work(0);
            for (auto &t : th)
                t.join();
            const auto t1 = std::chrono::steady_clock::now();
            for (int k = 0; k < n_ctx && st == MIA_HIP_OK; ++k)
                if (rcs[k] != MIA_HIP_OK)
                    st = fail(rcs[k], "optimize_many, context " + std::to_string(k) + ": " + msgs[k]);
            if (st == MIA_HIP_OK && seconds)
                *seconds = std::chrono::duration<double>(t1 - t0).count();
        }
        for (int k = 0; k < n_ctx && st == MIA_HIP_OK; ++k)
            st = opt_run_results(runs[k], x_out ? x_out[k] : nullptr, costs_out ? costs_out[k] : nullptr,
                                 evals_out ? evals_out[k] : nullptr);
    }
    return st;
}

int mia_hip_set_optimize_launch(mia_hip_ctx *c, int mode)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (mode != MIA_HIP_LAUNCH_STREAM && mode != MIA_HIP_LAUNCH_GRAPH)
        return fail(MIA_HIP_ERR_ARG, "optimize launch: mode must be MIA_HIP_LAUNCH_STREAM or MIA_HIP_LAUNCH_GRAPH");
    c->opt_launch = mode;
    return MIA_HIP_OK;
}

int mia_hip_synchronize(mia_hip_ctx *c)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MIA_HIP_OK;
}

int mia_hip_get_stream(mia_hip_ctx *c, void **stream)
{
    if (!c || !stream)
        return fail(MIA_HIP_ERR_ARG, "null pointer");
    *stream = (void *)c->stream;
    return MIA_HIP_OK;
}

int mia_hip_profile_enable(mia_hip_ctx *c, int enable)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    c->profile = enable != 0;
    return MIA_HIP_OK;
}

int mia_hip_profile_read(mia_hip_ctx *c, double *ms, int *launches, int max_stages, int *n_stages)
{
    if (!c)
        return fail(MIA_HIP_ERR_ARG, "null context");
    int st = collect_profile(c);
    if (st)
        return st;
    const int n = std::min<int>(max_stages, ST_COUNT);
    for (int k = 0; k < n; ++k) {
        if (ms)
            ms[k] = c->prof_ms[k];
        if (launches)
            launches[k] = c->prof_n[k];
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    if (n_stages)
        *n_stages = ST_COUNT;
    return MIA_HIP_OK;
}

const char *mia_hip_stage_name(int stage)
{
    if (stage < 0 || stage >= ST_COUNT)
        return "";
    return kStageNames[stage];
}

} // extern "C"

// ---------------------------------------------------------------------------
// tile variants, stage-1 layouts and their autotuner
// ---------------------------------------------------------------------------

int mia_hip_device_arch(int device, char *buf, int len)
{
    if (!buf || len <= 0)
        return fail(MIA_HIP_ERR_ARG, "null buffer");
    const std::string a = device_arch(device);
    if (a.empty())
        return fail(MIA_HIP_ERR_DEVICE, "cannot query device " + std::to_string(device));
    std::strncpy(buf, a.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return MIA_HIP_OK;
}

int mia_hip_warp_variants(int *tile_dims, int max_variants, int *n_variants)
{
    if (n_variants)
        *n_variants = kW3Variants;
    for (int v = 0; v < kW3Variants && v < max_variants && tile_dims; ++v)
        w3_tile_dims(v, tile_dims[3 * v], tile_dims[3 * v + 1], tile_dims[3 * v + 2]);
    return MIA_HIP_OK;
}

int mia_hip_warp2d_variants(int *tile_dims, int max_variants, int *n_variants)
{
    if (n_variants)
        *n_variants = kW2Variants;
    for (int v = 0; v < kW2Variants && v < max_variants && tile_dims; ++v) {
        tile_dims[3 * v] = w2_tile_x(v);
        tile_dims[3 * v + 1] = w2_tile_y(v);
        tile_dims[3 * v + 2] = w2_box(v) ? 1 : 0;
    }
    return MIA_HIP_OK;
}

int mia_hip_set_warp_variant(mia_hip_ctx *ctx, int variant)
{
    if (!ctx)
        return fail(MIA_HIP_ERR_ARG, "null context");
    const int nv = std::max(kW3Variants, kW2Variants);
    if (variant < -1 || variant >= nv)
        return fail(MIA_HIP_ERR_ARG, "warp variant must be -1 (tuned) or 0.." + std::to_string(nv - 1) +
                                         " (3D: 0.." + std::to_string(kW3Variants - 1) + ")");
    ctx->force_variant = variant;
    return MIA_HIP_OK;
}

int mia_hip_get_warp_variant(mia_hip_ctx *ctx, int inst, int *variant)
{
    if (!ctx || !variant)
        return fail(MIA_HIP_ERR_ARG, "null argument");
    if (ctx->n_inst == 0)
        return fail(MIA_HIP_ERR_STATE, "mia_hip_setup has not been called");
    if (inst < 0 || inst >= ctx->n_inst)
        return fail(MIA_HIP_ERR_ARG, "instance index out of range");
    *variant = ctx->inst_variant[inst];
    return MIA_HIP_OK;
}

int mia_hip_set_stage1_path(mia_hip_ctx *ctx, int path)
{
    if (!ctx)
        return fail(MIA_HIP_ERR_ARG, "null context");
    if (path < -1 || path > 1)
        return fail(MIA_HIP_ERR_ARG, "stage-1 layout must be -1 (tuned), 0 (per slice) or 1 (per z knot cell)");
    ctx->force_s1 = path;
    return MIA_HIP_OK;
}

int mia_hip_load_warp_tuning(int device, const char *json_path)
{
    Tuning &t = tuning();
    std::lock_guard<std::mutex> lock(t.m);
    load_tuning_locked(t, device, json_path);
    if (json_path && t.source.empty())
        return fail(MIA_HIP_ERR_ARG, std::string("no tuning table for this GPU in ") + json_path);
    return MIA_HIP_OK;
}

int mia_hip_warp_tuning_source(char *buf, int len)
{
    Tuning &t = tuning();
    std::lock_guard<std::mutex> lock(t.m);
    if (buf && len > 0) {
        std::strncpy(buf, t.source.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)t.table.size();
}

namespace {
// deterministic N(0, sd) samples clipped to +-clip (Box-Muller on a 64-bit LCG)
std::vector<double> tuning_params(size_t n, double sd, double clip, unsigned long long seed)
{
    std::vector<double> p(n);
    unsigned long long x = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (size_t k = 0; k < n; ++k) {
        const double u1 = uni(), u2 = uni();
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        p[k] = std::max(-clip, std::min(clip, sd * z));
    }
    return p;
}

// the synthetic workload of one tuning measurement
struct TuneCase {
    int dim, n_inst, n; // n: side of the cube / square
    float rate;
};

TuneCase tune_case(int kind, int bucket)
{
    if (kind == MIA_HIP_TUNE_WARP2D)
        return {2, 16, (int)std::llround(std::pow(2.0, bucket / 2.0)), 8.0f};
    return {3, 1, (int)std::llround(std::pow(2.0, bucket / 3.0)), kind == MIA_HIP_TUNE_WARP3D ? 6.0f : 8.0f};
}

// mean microseconds per evaluation of the kind's stages with choice v forced
int time_choice(int device, int kind, int degree, const TuneCase &tc, int bucket, int v, double &us)
{
    uint32_t sz[3], cs[3];
    std::vector<uint32_t> sizes, csizes;
    const int c1 = mia_hip_spline_csize((uint32_t)tc.n, tc.rate, 3);
    for (int a = 0; a < 3; ++a) {
        sz[a] = (uint32_t)tc.n;
        cs[a] = (uint32_t)c1;
    }
    for (int k = 0; k < tc.n_inst; ++k) {
        sizes.insert(sizes.end(), sz, sz + 3);
        csizes.insert(csizes.end(), cs, cs + 3);
    }
    const size_t np = (size_t)tc.n_inst * tc.dim * c1 * c1 * (tc.dim == 3 ? c1 : 1);
    const std::vector<double> p = tuning_params(np, 1.5, 4.0, 1000 + bucket);
    std::vector<float> img((size_t)tc.n * tc.n * (tc.dim == 3 ? tc.n : 1), 0.0f);
    std::vector<double> g(np), cost(tc.n_inst);
    mia_hip_ctx *ctx = nullptr;
    int st = mia_hip_ctx_create(device, &ctx);
    if (st)
        return st;
    if (kind == MIA_HIP_TUNE_STAGE1)
        st = mia_hip_set_stage1_path(ctx, v);
    else
        st = mia_hip_set_warp_variant(ctx, v);
    if (!st)
        st = mia_hip_setup(ctx, tc.dim, tc.n_inst, sizes.data(), csizes.data(), degree);
    for (int k = 0; k < tc.n_inst && !st; ++k)
        st = mia_hip_set_images(ctx, k, img.data(), img.data());
    for (int k = 0; k < 2 && !st; ++k)
        st = mia_hip_cost_grad(ctx, p.data(), g.data(), cost.data());
    if (!st)
        st = mia_hip_profile_enable(ctx, 1);
    const int reps = 5;
    for (int k = 0; k < reps && !st; ++k)
        st = mia_hip_cost_grad(ctx, p.data(), g.data(), cost.data());
    double ms[ST_COUNT];
    int nl[ST_COUNT], ns = 0;
    if (!st)
        st = mia_hip_profile_read(ctx, ms, nl, ST_COUNT, &ns);
    mia_hip_ctx_destroy(ctx);
    if (st)
        return st;
    if (kind == MIA_HIP_TUNE_STAGE1)
        us = 1e3 * (ms[ST_STAGE1] + ms[ST_REDUCE]) / reps;
    else
        us = nl[ST_WARP] ? 1e3 * ms[ST_WARP] / nl[ST_WARP] : 1e300;
    return MIA_HIP_OK;
}
} // namespace

int mia_hip_autotune(int device, int kind, int degree, int b_min, int b_max, const char *json_path, double *us_out)
{
    if (kind < 0 || kind >= kTuneKinds)
        return fail(MIA_HIP_ERR_ARG, "autotune: unknown kind " + std::to_string(kind));
    if (degree != 0 && degree != 1 && degree != 3)
        return fail(MIA_HIP_ERR_ARG, "autotune: degree must be 1, 3 or 0 (both)");
    const int lo[kTuneKinds] = {12, 10, 12}, hi[kTuneKinds] = {26, 24, 24};
    if (b_min < lo[kind] || b_max > hi[kind] || b_min > b_max)
        return fail(MIA_HIP_ERR_ARG, std::string("autotune: size buckets of ") + kTuneKindNames[kind] +
                                         " must satisfy " + std::to_string(lo[kind]) + " <= b_min <= b_max <= " +
                                         std::to_string(hi[kind]));
    const std::string arch = device_arch(device);
    if (arch.empty())
        return fail(MIA_HIP_ERR_DEVICE, "autotune: cannot query device " + std::to_string(device));
    const int nc = kTuneChoices[kind];
    const int degs[2] = {3, 1};
    struct Row {
        int degree, bucket, variant;
        TuneCase tc;
        double us[kMaxTuneChoices];
    };
    std::vector<Row> rows;
    int k_out = 0;
    for (int di = 0; di < 2; ++di) {
        const int deg = degs[di];
        if (degree != 0 && degree != deg)
            continue;
        for (int b = b_min; b <= b_max; ++b) {
            Row r{deg, b, 0, tune_case(kind, b), {0}};
            double best = 1e300;
            for (int v = 0; v < nc; ++v) {
                const int st = time_choice(device, kind, deg, r.tc, b, v, r.us[v]);
                if (st)
                    return st;
                if (r.us[v] < best) {
                    best = r.us[v];
                    r.variant = v;
                }
                if (us_out)
                    us_out[k_out * nc + v] = r.us[v];
            }
            ++k_out;
            rows.push_back(r);
        }
    }
    if (!json_path)
        return MIA_HIP_OK;
    // the entries an existing table of this architecture holds for other
    // kinds, degrees or buckets
    std::vector<std::string> keep;
    {
        std::string text, file_arch;
        std::map<std::tuple<int, int, int>, int> old;
        if (parse_tuning(json_path, file_arch, old) && file_arch == arch && read_text(json_path, text)) {
            size_t pos = 0;
            while ((pos = text.find("\"bucket\"", pos)) != std::string::npos) {
                const size_t b = text.rfind('{', pos), e = text.find('}', pos);
                ++pos;
                if (b == std::string::npos || e == std::string::npos)
                    continue;
                int d = 0, bk = 0;
                const bool retuned = tune_kind(json_string(text, b, e, "kernel")) == kind &&
                                     json_int(text, b, e, "degree", d) && (degree == 0 || d == degree) &&
                                     json_int(text, b, e, "bucket", bk) && bk >= b_min && bk <= b_max;
                if (!retuned)
                    keep.push_back(text.substr(b, e - b + 1));
            }
        }
    }
    std::FILE *f = std::fopen(json_path, "w");
    if (!f)
        return fail(MIA_HIP_ERR_ARG, std::string("autotune: cannot write ") + json_path);
    std::fprintf(f, "{\"arch\": \"%s\",\n \"variants\": {\"warp3d\": [", arch.c_str());
    for (int v = 0; v < kW3Variants; ++v) {
        int tx, ty, tz;
        w3_tile_dims(v, tx, ty, tz);
        std::fprintf(f, "%s[%d, %d, %d]", v ? ", " : "", tx, ty, tz);
    }
    std::fprintf(f, "], \"warp2d\": [");
    for (int v = 0; v < kW2Variants; ++v)
        std::fprintf(f, "%s[%d, %d, \"%s\"]", v ? ", " : "", w2_tile_x(v), w2_tile_y(v), w2_box(v) ? "box" : "direct");
    std::fprintf(f, "], \"stage1_3d\": [\"slice\", \"zcell\"]},\n \"table\": [\n");
    size_t left = keep.size() + rows.size();
    for (const std::string &e : keep)
        std::fprintf(f, "  %s%s\n", e.c_str(), --left ? "," : "");
    for (const Row &r : rows) {
        std::fprintf(f, "  {\"kernel\": \"%s\", \"degree\": %d, \"bucket\": %d, \"variant\": %d, \"instances\": %d, "
                        "\"size\": [%d, %d%s], \"rate\": %g, \"us\": [",
                     kTuneKindNames[kind], r.degree, r.bucket, r.variant, r.tc.n_inst, r.tc.n, r.tc.n,
                     r.tc.dim == 3 ? (", " + std::to_string(r.tc.n)).c_str() : "", (double)r.tc.rate);
        for (int v = 0; v < nc; ++v)
            std::fprintf(f, "%s%.2f", v ? ", " : "", r.us[v]);
        std::fprintf(f, "]}%s\n", --left ? "," : "");
    }
    std::fprintf(f, " ]}\n");
    std::fclose(f);
    return mia_hip_load_warp_tuning(device, json_path);
}

int mia_hip_autotune_warp(int device, int degree, int b_min, int b_max, const char *json_path, double *us_out)
{
    return mia_hip_autotune(device, MIA_HIP_TUNE_WARP3D, degree, b_min, b_max, json_path, us_out);
}

// ---------------------------------------------------------------------------
// allocation-free evaluation of one instance
// ---------------------------------------------------------------------------

namespace {

// Scratch memory of mia_hip_cost_grad_noalloc: one array per GPU, held until
// the process ends.  In host code the array's name is a stand-in, not a
// device address: each device's address comes from hipGetSymbolAddress, taken
// once with that device current and kept per device; host code passes only
// that address on.
static __device__ __align__(256) unsigned char scratch[MIA_HIP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;

struct ScratchAddr {
    std::mutex m;
    unsigned char *ptr[kMaxScratchDevices] = {};
};

ScratchAddr &scratch_addr()
{
    static ScratchAddr s;
    return s;
}

// the device's scratch address (device current, 0 <= device < kMaxScratchDevices)
int scratch_ptr(int device, unsigned char **out)
{
    ScratchAddr &s = scratch_addr();
    std::lock_guard<std::mutex> lock(s.m);
    if (!s.ptr[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return fail(MIA_HIP_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch memory");
        s.ptr[device] = (unsigned char *)p;
    }
    *out = s.ptr[device];
    return MIA_HIP_OK;
}

} // namespace

int mia_hip_cost_grad_noalloc(int device, int dim, const uint32_t *size, const uint32_t *csize, int degree, int norm,
                              float autothresh, double weight, int grad_mode, const float *d_moving,
                              const float *d_fixed, const double *d_params, double *d_grad, double *d_cost,
                              void *stream)
{
    if (dim != 2 && dim != 3)
        return fail(MIA_HIP_ERR_ARG, "dim must be 2 or 3");
    if (!size || !csize || !d_moving || !d_fixed || !d_params || !d_grad || !d_cost)
        return fail(MIA_HIP_ERR_ARG, "null pointer");
    if (degree != 1 && degree != 3)
        return fail(MIA_HIP_ERR_ARG, "image interpolation kernel must be bspline:d=1 or bspline:d=3");
    if (!(autothresh >= 0.0f && autothresh <= 1000.0f))
        return fail(MIA_HIP_ERR_ARG, "autothresh must be in [0, 1000]");
    if (grad_mode != MIA_HIP_GRAD_REFERENCE && grad_mode != MIA_HIP_GRAD_EXACT)
        return fail(MIA_HIP_ERR_ARG, "unknown gradient mode");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count || device >= kMaxScratchDevices)
        return fail(MIA_HIP_ERR_ARG, "no HIP device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    unsigned char *base = nullptr;
    int st = scratch_ptr(device, &base);
    if (st)
        return st;

    // the instance as mia_hip_setup describes the only instance of a batch
    InstDesc d{};
    d.nx = (int)size[0];
    d.ny = (int)size[1];
    d.nz = dim == 3 ? (int)size[2] : 1;
    d.cx = (int)csize[0];
    d.cy = (int)csize[1];
    d.cz = dim == 3 ? (int)csize[2] : 1;
    if (d.nx > kMaxRow)
        return fail(MIA_HIP_ERR_LIMIT,
                    "image x size " + std::to_string(d.nx) + " exceeds the limit " + std::to_string(kMaxRow));
    const long long nvox = (long long)d.nx * d.ny * d.nz;
    if (dim == 3 && nvox * (long long)sizeof(float) >= (1LL << 31))
        return fail(MIA_HIP_ERR_LIMIT, "3D image has 2^29 voxels or more (limit of the 32-bit image offsets)");
    AxisCheck ax, ay, az;
    if ((st = check_axis(d.nx, d.cx, "x", ax)) || (st = check_axis(d.ny, d.cy, "y", ay)))
        return st;
    if (dim == 3) {
        if ((st = check_axis(d.nz, d.cz, "z", az)))
            return st;
    } else {
        az.nbands = 1;
        az.band_rows = 1;
    }
    d.ncx = ax.nbands;
    d.ncy = ay.nbands;
    d.ncz = az.nbands;
    d.zcell = zcell_eligible(dim, d.nx, ay.band_rows, (long long)d.cx * d.cy * d.cz) &&
                      tuned_choice(device, MIA_HIP_TUNE_STAGE1, degree, nvox) == 1
                  ? 1
                  : 0;
    d.img_off = d.par_off = d.part_off = 0;
    d.tab_x = 0;
    d.tab_y = d.nx;
    d.tab_z = d.nx + d.ny;
    d.cell_x = 0;
    d.cell_y = d.cx + 1;
    d.cell_z = d.cx + 1 + d.cy + 1;
    const long long ncoef = (long long)d.cx * d.cy * d.cz;
    const long long part = dim == 3 ? std::max((long long)d.nz * d.ncy * 12 * d.cx, (long long)d.ncz * d.ncy * 48 * d.cx)
                                    : (long long)d.ncy * 8 * d.cx;
    OneInstCounts nc;
    nc.s1 = dim == 3 ? d.nz * d.ncy : d.ncy;
    nc.red = (int)((dim * ncoef + (dim == 3 ? 64 : kBlock) - 1) / (dim == 3 ? 64 : kBlock));
    nc.s1z = d.ncz * d.ncy;
    nc.redz = (int)((dim * ncoef + kBlock - 1) / kBlock);
    const int variant = tuned_choice(device, dim == 3 ? MIA_HIP_TUNE_WARP3D : MIA_HIP_TUNE_WARP2D, degree, nvox);
    int w2_win = 0;
    if (dim == 3) {
        int tx, ty, tz;
        w3_tile_dims(variant, tx, ty, tz);
        nc.tiles = ((d.nx + tx - 1) / tx) * ((d.ny + ty - 1) / ty) * ((d.nz + tz - 1) / tz);
    } else {
        const int tx = w2_tile_x(variant), ty = w2_tile_y(variant);
        nc.tiles = ((d.nx + tx - 1) / tx) * ((d.ny + ty - 1) / ty);
        // the coefficient window of a tile (from the same start indices as
        // the context's axis tables)
        float w[4];
        w2_win = tile_span(d.nx, tx, [&](int p) { return spline_axis_entry(p, d.nx, d.cx, w); }) *
                 tile_span(d.ny, ty, [&](int p) { return spline_axis_entry(p, d.ny, d.cy, w); });
        if (w2_win > kW2Win)
            return fail(MIA_HIP_ERR_ARG, "2D coefficient window exceeds its bound (spline rate below 1?)");
    }

    // scratch layout (256-byte aligned pieces); nothing in it needs to start
    // at zero: every element is written before it is read in each call
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return base + o;
    };
    const int ntab = d.nx + d.ny + d.nz, ncell = d.cx + d.cy + d.cz + 3;
    InstDesc *s_inst = (InstDesc *)take(sizeof(InstDesc));
    int *s_ints = (int *)take(sizeof(int) * kOneInstInts);
    float4 *s_tw = (float4 *)take(sizeof(float4) * ntab);
    int *s_ts = (int *)take(sizeof(int) * ntab);
    int *s_cells = (int *)take(sizeof(int) * ncell);
    float2 *s_clamp = (float2 *)take(sizeof(float2));
    float2 *s_mm = (float2 *)take(sizeof(float2) * 1024);
    double *s_gscale = (double *)take(sizeof(double));
    double *s_ssd = (double *)take(sizeof(double) * (nc.s1 + nc.s1z));
    unsigned *s_cnt = (unsigned *)take(sizeof(unsigned) * (nc.s1 + nc.s1z));
    float *s_M = (float *)take(sizeof(float) * nvox);
    float *s_W = (float *)take(sizeof(float) * nvox);
    float *s_P = (float *)take(sizeof(float) * part);
    double *s_pf = degree == 3 ? (double *)take(sizeof(double) * nvox) : nullptr;
    if (off > (size_t)MIA_HIP_NOALLOC_SCRATCH_BYTES)
        return fail(MIA_HIP_ERR_LIMIT, "allocation-free evaluation needs " + std::to_string(off) +
                                           " bytes of scratch memory, more than its " +
                                           std::to_string((size_t)MIA_HIP_NOALLOC_SCRATCH_BYTES));
    if ((st = set_stage1_lds(dim, d.nx, d.cx, d.ncx)))
        return st;

    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_one_inst_setup, dim3(1), dim3(kBlock), 0, s, d, dim, nc, s_inst, s_ints, s_tw, s_ts,
                       s_cells);
    LAUNCH_CHECK("one-instance set-up");
    HIP_TRY(hipMemcpyAsync(s_M, d_moving, sizeof(float) * nvox, hipMemcpyDeviceToDevice, s));
    if ((st = enqueue_prefilter(d, dim, degree, s_M, s_mm, s_clamp, s_pf, s)))
        return st;

    EvalPlan pl;
    pl.dim = dim;
    pl.degree = degree;
    pl.n_inst = 1;
    pl.n_zcell = d.zcell;
    pl.norm = norm ? 1 : 0;
    pl.grad_mode = grad_mode;
    pl.thr = autothresh;
    pl.weight = weight;
    pl.max_nx = d.nx;
    pl.max_cx = d.cx;
    pl.max_ncx = d.ncx;
    pl.w2_win = w2_win;
    pl.s1_blocks = nc.s1;
    pl.red_blocks = nc.red;
    pl.s1z_blocks = nc.s1z;
    pl.redz_blocks = nc.redz;
    pl.n_warp = 1;
    pl.warp[0].variant = variant;
    pl.warp[0].n = 1;
    pl.warp[0].tiles = nc.tiles;
    pl.warp[0].ids = s_ints + 8;
    pl.warp[0].prefix = s_ints + 9;
    pl.inst = s_inst;
    pl.s1_prefix = s_ints;
    pl.red_prefix = s_ints + 2;
    pl.s1z_prefix = s_ints + 4;
    pl.redz_prefix = s_ints + 6;
    pl.M = s_M;
    pl.R = d_fixed;
    pl.W = s_W;
    pl.P = s_P;
    pl.tw = s_tw;
    pl.ts = s_ts;
    pl.cells = s_cells;
    pl.clamp = s_clamp;
    pl.ssd = s_ssd;
    pl.gscale = s_gscale;
    pl.cnt = s_cnt;
    return enqueue_plan(pl, s, nullptr, d_params, d_grad, d_cost);
}
