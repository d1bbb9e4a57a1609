// api.cpp -- the extern "C" boundary (include/qslhip.h).
//
// Every entry point validates its arguments on the host before any device
// work, maps the reference's std::logic_error cases to QSLHIP_ESIZE /
// QSLHIP_ENOTNP / QSLHIP_EINVAL, and turns HIP failures into
// QSLHIP_ENOMEM / QSLHIP_ENODEV / QSLHIP_EDEVICE with the detail kept in a
// thread-local string (qslhip_last_error).  No exception crosses the ABI.
#include <cmath>
#include <cstring>
#include <random>

#include "dense.h"
#include "rtile.h"
#include "gate_math.h"

using namespace qslhip;

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F &&f)
{
    try {
        g_last_error.clear();
        f();
        return QSLHIP_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return QSLHIP_ENOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return QSLHIP_EDEVICE;
    } catch (...) {
        g_last_error = "unknown failure";
        return QSLHIP_EDEVICE;
    }
}

void require_handle(qslhip_t h)
{
    QH_REQUIRE(h != nullptr && h->amps != nullptr, QSLHIP_EINVAL, "null handle");
}

int first_device()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1)
        throw Error(QSLHIP_ENODEV, std::string("no HIP device: ") +
                                       (e != hipSuccess ? hipGetErrorString(e)
                                                        : "device count is 0"));
    int dev = 0;
    QH_CHECK(hipGetDevice(&dev));
    return dev;
}

uint64_t choose_u64(unsigned n, unsigned k)
{
    if (k > n) return 0;
    uint64_t a = 1;
    for (unsigned i = 1; i <= k; i++) {
        a *= (uint64_t)(n - k + i);
        a /= i;
    }
    return a;
}

qslhip_state *alloc_handle(unsigned n, int prec, bool np, unsigned k)
{
    int dev = first_device();
    QH_CHECK(hipSetDevice(dev));
    auto *h = new qslhip_state();
    h->device = dev;
    h->n = n;
    h->prec = prec;
    h->np = np;
    h->k = k;
    h->namps = np ? choose_u64(n, k) : (1ull << n);
    h->cap = h->namps;
    try {
        QH_CHECK(hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking));
        h->stream = h->own_stream;
        QH_CHECK(hipEventCreateWithFlags(&h->staged, hipEventDisableTiming));
        size_t bytes = h->namps * h->amp_bytes();
        hipError_t e = hipMalloc(&h->amps, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            h->amps = nullptr;
            throw Error(QSLHIP_ENOMEM, "hipMalloc of " + std::to_string(bytes) +
                                           " bytes for the state failed: " +
                                           hipGetErrorString(e));
        }
        if (np) np_init(h);
    } catch (...) {
        qslhip_destroy(h);
        throw;
    }
    return h;
}

void reset_impl(qslhip_t h)
{
    // Default: |0...0> (default.cpp:57-65), recorded and written by the next
    // reader (materialize) or synthesized by the next gate pass.  NP: the
    // lowest weight-k string 2^k - 1, whose colex rank is 0 (np.cpp:68-79).
    if (h->np) launch_fill_basis(h, 0);
    else h->pending_basis = 0;
    h->sums_valid = false;
}

// Writes a recorded basis state into the vector, before anything reads it.
void materialize(qslhip_t h)
{
    if (h->pending_basis == qslhip_state::kNoBasis) return;
    launch_fill_basis(h, h->pending_basis);
    h->pending_basis = qslhip_state::kNoBasis;
}

// A handle whose vector holds its state, for an entry point that only
// reads it.
void require_state_ro(qslhip_t h)
{
    require_handle(h);
    materialize(h);
}

// The same for one that changes the state: the sampling sums a gate pass
// left no longer describe it.
void require_state(qslhip_t h)
{
    require_state_ro(h);
    h->sums_valid = false;
}

template <typename Fp>
void check_np_host(const void *host, uint64_t count, unsigned *nones)
{
    // checkStateNP (src/utils/quantum.cpp:133-156): every nonzero
    // amplitude must carry the same Hamming weight.
    const Fp *p = static_cast<const Fp *>(host);
    bool found = false;
    unsigned w0 = 0;
    for (uint64_t i = 0; i < count; i++) {
        Fp a = p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1];
        if (a != 0) {
            unsigned w = (unsigned)__builtin_popcountll(i);
            if (!found) {
                w0 = w;
                found = true;
            } else if (w != w0) {
                throw Error(QSLHIP_ENOTNP, "Input state is not number preserving.");
            }
        }
    }
    *nones = w0;
}

}  // namespace

namespace qslhip {

int guarded_call(const std::function<void()> &f) { return guarded(f); }

void DevBuf::ensure(size_t b)
{
    if (b <= bytes) return;
    release();
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        bytes = 0;
        throw Error(QSLHIP_ENOMEM, "hipMalloc of " + std::to_string(b) +
                                       " scratch bytes failed");
    }
    bytes = b;
}

void DevBuf::release()
{
    if (ptr) {
        // frees wait for every queued kernel that may still read the buffer
        (void)hipDeviceSynchronize();
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
}

void PinnedBuf::ensure(size_t b)
{
    if (b <= bytes) return;
    release();
    QH_CHECK(hipHostMalloc(&ptr, b, hipHostMallocDefault));
    bytes = b;
}

void PinnedBuf::release()
{
    if (ptr) {
        (void)hipDeviceSynchronize();
        (void)hipHostFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
}

Tuning &tuning()
{
    static Tuning t;
    return t;
}

}  // namespace qslhip

extern "C" {

const char *qslhip_status_string(int status)
{
    switch (status) {
    case QSLHIP_OK: return "ok";
    case QSLHIP_EINVAL: return "invalid argument";
    case QSLHIP_ESIZE: return "state size mismatch";
    case QSLHIP_ENOTNP: return "state is not number preserving";
    case QSLHIP_ENOMEM: return "device out of memory";
    case QSLHIP_ENODEV: return "no HIP device";
    case QSLHIP_EDEVICE: return "HIP runtime error";
    case QSLHIP_EKIND: return "gate not available on this simulator type";
    default: return "unknown status";
    }
}

const char *qslhip_last_error(void) { return g_last_error.c_str(); }

int qslhip_device_count(int *count)
{
    return guarded([&] {
        QH_REQUIRE(count, QSLHIP_EINVAL, "null count");
        *count = 0;
        hipError_t e = hipGetDeviceCount(count);
        if (e != hipSuccess) {
            *count = 0;
            throw Error(QSLHIP_ENODEV, hipGetErrorString(e));
        }
        if (*count < 1) throw Error(QSLHIP_ENODEV, "device count is 0");
    });
}

int qslhip_create(unsigned nqubits, int prec, qslhip_t *out)
{
    return guarded([&] {
        QH_REQUIRE(out, QSLHIP_EINVAL, "null output handle");
        *out = nullptr;
        QH_REQUIRE(prec == 4 || prec == 8, QSLHIP_EINVAL, "prec must be 4 or 8");
        QH_REQUIRE(nqubits >= 1 && nqubits <= kMaxQubits, QSLHIP_EINVAL,
                   "nqubits must be in [1, 40]");
        qslhip_t h = alloc_handle(nqubits, prec, false, 0);
        reset_impl(h);
        QH_CHECK(hipStreamSynchronize(h->stream));
        *out = h;
    });
}

int qslhip_create_resize(unsigned nqubits, int prec, qslhip_t *out)
{
    return guarded([&] {
        QH_REQUIRE(out, QSLHIP_EINVAL, "null output handle");
        *out = nullptr;
        QH_REQUIRE(prec == 4 || prec == 8, QSLHIP_EINVAL, "prec must be 4 or 8");
        QH_REQUIRE(nqubits >= 1 && nqubits <= kMaxQubits, QSLHIP_EINVAL,
                   "nqubits must be in [1, 40]");
        qslhip_t h = alloc_handle(nqubits, prec, false, 0);
        h->resizable = true;
        reset_impl(h);
        QH_CHECK(hipStreamSynchronize(h->stream));
        *out = h;
    });
}

int qslhip_create_np(unsigned nqubits, unsigned nones, int prec, qslhip_t *out)
{
    return guarded([&] {
        QH_REQUIRE(out, QSLHIP_EINVAL, "null output handle");
        *out = nullptr;
        QH_REQUIRE(prec == 4 || prec == 8, QSLHIP_EINVAL, "prec must be 4 or 8");
        QH_REQUIRE(nqubits >= 2 && nqubits <= 62, QSLHIP_EINVAL,
                   "nqubits must be in [2, 62]");
        // np.cpp:48-50
        QH_REQUIRE(nones >= 1 && nones < nqubits, QSLHIP_ENOTNP,
                   "The number of ones must be 1 <= nones < nqubits.");
        QH_REQUIRE(choose_u64(nqubits, nones) <= (1ull << 36), QSLHIP_EINVAL,
                   "C(n, nones) too large for one device");
        qslhip_t h = alloc_handle(nqubits, prec, true, nones);
        reset_impl(h);
        QH_CHECK(hipStreamSynchronize(h->stream));
        *out = h;
    });
}

int qslhip_destroy(qslhip_t h)
{
    return guarded([&] {
        if (!h) return;
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        if (h->np) np_free(h);
        h->red.release();
        h->cdf.release();
        h->shots.release();
        h->optab.release();
        h->pin.release();
        if (h->amps) (void)hipFree(h->amps);
        if (h->staged) (void)hipEventDestroy(h->staged);
        if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
        for (hipEvent_t e : h->xev)
            if (e) (void)hipEventDestroy(e);
        if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
        if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
        delete h;
    });
}

int qslhip_clone(qslhip_t src, qslhip_t *out)
{
    return guarded([&] {
        require_state_ro(src);
        QH_REQUIRE(out, QSLHIP_EINVAL, "null output handle");
        *out = nullptr;
        qslhip_t h = alloc_handle(src->n, src->prec, src->np, src->k);
        h->resizable = src->resizable;
        QH_CHECK(hipStreamSynchronize(src->stream));
        QH_CHECK(hipMemcpyAsync(h->amps, src->amps, src->namps * src->amp_bytes(),
                                hipMemcpyDeviceToDevice, h->stream));
        QH_CHECK(hipStreamSynchronize(h->stream));
        *out = h;
    });
}

int qslhip_assign(qslhip_t dst, qslhip_t src)
{
    return guarded([&] {
        require_handle(dst);
        require_state_ro(src);
        if (dst == src) return;
        // default.cpp:94-105 / np.cpp:119-131.  Every check comes before
        // any change to dst, so a refused assignment leaves it as it was
        // (a recorded basis state stays recorded).
        QH_REQUIRE(dst->n == src->n, QSLHIP_ESIZE,
                   "Cannot assign Qubits object vector from different number of qubits");
        QH_REQUIRE(dst->np == src->np && dst->prec == src->prec, QSLHIP_EINVAL,
                   "handles differ in type or precision");
        dst->pending_basis = qslhip_state::kNoBasis;  // overwritten below
        dst->sums_valid = false;
        if (dst->np && dst->k != src->k) {
            // the NP reference copies the full vector, whatever its weight
            np_free(dst);
            (void)hipFree(dst->amps);
            dst->amps = nullptr;
            dst->k = src->k;
            dst->namps = src->namps;
            QH_CHECK(hipMalloc(&dst->amps, dst->namps * dst->amp_bytes()));
            np_init(dst);
        }
        QH_CHECK(hipStreamSynchronize(src->stream));
        QH_CHECK(hipMemcpyAsync(dst->amps, src->amps, src->namps * src->amp_bytes(),
                                hipMemcpyDeviceToDevice, dst->stream));
        QH_CHECK(hipStreamSynchronize(dst->stream));
    });
}

int qslhip_info(qslhip_t h, unsigned *nqubits, unsigned *nones, int *prec,
                uint64_t *namps)
{
    return guarded([&] {
        require_handle(h);
        if (nqubits) *nqubits = h->n;
        if (nones) *nones = h->np ? h->k : 0;
        if (prec) *prec = h->prec;
        if (namps) *namps = h->namps;
    });
}

int qslhip_set_stream(qslhip_t h, void *stream)
{
    return guarded([&] {
        require_handle(h);
        QH_CHECK(hipStreamSynchronize(h->stream));
        h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
    });
}

int qslhip_get_stream(qslhip_t h, void **stream)
{
    return guarded([&] {
        require_handle(h);
        QH_REQUIRE(stream, QSLHIP_EINVAL, "null output");
        *stream = static_cast<void *>(h->stream);
    });
}

int qslhip_sync(qslhip_t h)
{
    return guarded([&] {
        require_handle(h);
        QH_CHECK(hipStreamSynchronize(h->stream));
    });
}

int qslhip_reset(qslhip_t h)
{
    return guarded([&] {
        require_handle(h);
        reset_impl(h);
    });
}

int qslhip_set_state(qslhip_t h, const void *host, uint64_t count)
{
    return guarded([&] {
        require_handle(h);
        QH_REQUIRE(host, QSLHIP_EINVAL, "null host pointer");
        // default.cpp:70-74
        QH_REQUIRE(count == (1ull << h->n), QSLHIP_ESIZE,
                   "Cannot assign state vector from different number of qubits");
        h->sums_valid = false;
        if (!h->np) {
            h->pending_basis = qslhip_state::kNoBasis;  // overwritten
            QH_CHECK(hipMemcpyAsync(h->amps, host, count * h->amp_bytes(),
                                    hipMemcpyHostToDevice, h->stream));
            QH_CHECK(hipStreamSynchronize(h->stream));
            return;
        }
        // NP: re-derive nones (np.cpp:81-98), then gather the weight-k
        // amplitudes into the compact layout on the device.
        unsigned k = 0;
        if (h->prec == 8) check_np_host<double>(host, count, &k);
        else check_np_host<float>(host, count, &k);
        QH_REQUIRE(k >= 1 && k < h->n, QSLHIP_ENOTNP,
                   "The number of ones must be 1 <= nones < nqubits.");
        if (k != h->k) {
            np_free(h);
            (void)hipFree(h->amps);
            h->amps = nullptr;
            h->k = k;
            h->namps = choose_u64(h->n, k);
            QH_CHECK(hipMalloc(&h->amps, h->namps * h->amp_bytes()));
            np_init(h);
        }
        DevBuf full;
        full.ensure(count * h->amp_bytes());
        QH_CHECK(hipMemcpyAsync(full.ptr, host, count * h->amp_bytes(),
                                hipMemcpyHostToDevice, h->stream));
        np_gather_full(h, full.ptr);
        QH_CHECK(hipStreamSynchronize(h->stream));
        full.release();
    });
}

int qslhip_get_state(qslhip_t h, void *host, uint64_t count)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(host, QSLHIP_EINVAL, "null host pointer");
        QH_REQUIRE(count == (1ull << h->n), QSLHIP_ESIZE, "count must be 2^nqubits");
        if (!h->np) {
            QH_CHECK(hipMemcpyAsync(host, h->amps, count * h->amp_bytes(),
                                    hipMemcpyDeviceToHost, h->stream));
            QH_CHECK(hipStreamSynchronize(h->stream));
            return;
        }
        DevBuf full;
        full.ensure(count * h->amp_bytes());
        np_scatter_full(h, full.ptr);
        QH_CHECK(hipMemcpyAsync(host, full.ptr, count * h->amp_bytes(),
                                hipMemcpyDeviceToHost, h->stream));
        QH_CHECK(hipStreamSynchronize(h->stream));
        full.release();
    });
}

int qslhip_get_amps(qslhip_t h, void *host, uint64_t first, uint64_t count)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(host || count == 0, QSLHIP_EINVAL, "null host pointer");
        QH_REQUIRE(first <= h->namps && count <= h->namps - first, QSLHIP_EINVAL,
                   "amplitude range out of bounds");
        QH_CHECK(hipMemcpyAsync(host,
                                static_cast<char *>(h->amps) + first * h->amp_bytes(),
                                count * h->amp_bytes(), hipMemcpyDeviceToHost,
                                h->stream));
        QH_CHECK(hipStreamSynchronize(h->stream));
    });
}

int qslhip_set_amps(qslhip_t h, const void *host, uint64_t first, uint64_t count)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(host || count == 0, QSLHIP_EINVAL, "null host pointer");
        QH_REQUIRE(first <= h->namps && count <= h->namps - first, QSLHIP_EINVAL,
                   "amplitude range out of bounds");
        QH_CHECK(hipMemcpyAsync(static_cast<char *>(h->amps) + first * h->amp_bytes(),
                                host, count * h->amp_bytes(), hipMemcpyHostToDevice,
                                h->stream));
        QH_CHECK(hipStreamSynchronize(h->stream));
    });
}

int qslhip_set_basis(qslhip_t h, uint64_t index)
{
    return guarded([&] {
        require_handle(h);
        QH_REQUIRE(h->n >= 64 || index < (1ull << h->n), QSLHIP_EINVAL,
                   "basis index out of range");
        h->sums_valid = false;
        if (!h->np) {
            h->pending_basis = index;  // written by the next reader or gate pass
            return;
        }
        // np.cpp:100-116: nones follows the index's popcount
        unsigned k = (unsigned)__builtin_popcountll(index);
        QH_REQUIRE(k >= 1 && k < h->n, QSLHIP_ENOTNP,
                   "The number of ones must be 1 <= nones < nqubits.");
        if (k != h->k) {
            QH_CHECK(hipStreamSynchronize(h->stream));
            np_free(h);
            (void)hipFree(h->amps);
            h->amps = nullptr;
            h->k = k;
            h->namps = choose_u64(h->n, k);
            QH_CHECK(hipMalloc(&h->amps, h->namps * h->amp_bytes()));
            np_init(h);
        }
        launch_fill_basis(h, np_rank_host(index));
    });
}

int qslhip_set_num_ones(qslhip_t h, unsigned nones)
{
    return guarded([&] {
        require_handle(h);
        QH_REQUIRE(h->np, QSLHIP_EKIND, "setNumOnes needs a number-preserving handle");
        QH_REQUIRE(nones >= 1 && nones < h->n, QSLHIP_ENOTNP,
                   "The number of ones must be 1 <= nones < nqubits.");
        QH_REQUIRE(choose_u64(h->n, nones) <= (1ull << 36), QSLHIP_EINVAL,
                   "C(n, nones) too large for one device");
        if (nones != h->k) {
            QH_CHECK(hipStreamSynchronize(h->stream));
            np_free(h);
            (void)hipFree(h->amps);
            h->amps = nullptr;
            h->k = nones;
            h->namps = choose_u64(h->n, nones);
            QH_CHECK(hipMalloc(&h->amps, h->namps * h->amp_bytes()));
            np_init(h);
        }
        reset_impl(h);
    });
}

static void validate_gate(qslhip_t h, int kind, unsigned q0, unsigned q1)
{
    if (is_gate1(kind)) {
        QH_REQUIRE(q0 < h->n, QSLHIP_EINVAL, "qubit index out of range");
    } else if (is_gate2(kind)) {
        QH_REQUIRE(q0 < h->n && q1 < h->n, QSLHIP_EINVAL, "qubit index out of range");
        QH_REQUIRE(q0 != q1, QSLHIP_EINVAL, "two-qubit gate on one qubit");
    } else {
        throw Error(QSLHIP_EINVAL, "unknown gate kind " + std::to_string(kind));
    }
    if (h->np)
        QH_REQUIRE(is_np_kind(kind), QSLHIP_EKIND,
                   "gate does not preserve the number of ones");
}

int qslhip_gate1(qslhip_t h, int kind, unsigned targ, double angle)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(is_gate1(kind), QSLHIP_EINVAL, "not a one-qubit gate kind");
        validate_gate(h, kind, targ, 0);
        GateDesc g = make_desc(h->prec, kind, targ, 0, angle);
        if (h->np) np_launch_gate(h, g);
        else launch_gate(h, g);
    });
}

int qslhip_gate2(qslhip_t h, int kind, unsigned q0, unsigned q1, double angle)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(is_gate2(kind), QSLHIP_EINVAL, "not a two-qubit gate kind");
        validate_gate(h, kind, q0, q1);
        GateDesc g = make_desc(h->prec, kind, q0, q1, angle);
        if (h->np) np_launch_gate(h, g);
        else launch_gate(h, g);
    });
}

// Tile width of a fused pass on an n-qubit state: the tuned width, narrowed
// on small states so that at least 64 workgroups share the device, but not
// below 2^10 while register tiles are on (their narrowest kernel).
static int tile_width(const Tuning &tu, unsigned n, int prec)
{
    const int logt = prec == 8 ? tu.logt_f64 : tu.logt_f32;
    const int floor = tu.reg_tile ? 10 : 6;
    const int cap = (int)n - 6 > floor ? (int)n - 6 : floor;
    return logt < cap ? logt : cap;
}

int qslhip_apply_ops(qslhip_t h, const qslhip_op *ops, uint64_t nops, int mode)
{
    return guarded([&] {
        require_handle(h);
        QH_REQUIRE(ops || nops == 0, QSLHIP_EINVAL, "null op list");
        h->sums_valid = false;
        QH_REQUIRE(mode == QSLHIP_MODE_PER_GATE || mode == QSLHIP_MODE_FUSED ||
                       mode == QSLHIP_MODE_DENSE || mode == QSLHIP_MODE_REORDER,
                   QSLHIP_EINVAL, "unknown mode");
        if (mode == QSLHIP_MODE_DENSE) {
            QH_REQUIRE(!h->np && h->prec == 4, QSLHIP_EINVAL,
                       "dense MFMA mode needs a full-vector FP32 handle");
            QH_REQUIRE(h->n >= (unsigned)kDenseTile, QSLHIP_EINVAL,
                       "dense MFMA mode needs at least 12 qubits");
        }
        std::vector<GateDesc> gates;
        gates.reserve(nops);
        for (uint64_t i = 0; i < nops; i++) {
            const qslhip_op &o = ops[i];
            validate_gate(h, o.kind, o.q0, o.q1);
            gates.push_back(make_desc(h->prec, o.kind, o.q0,
                                      is_gate2(o.kind) ? o.q1 : 0, o.angle));
        }
        if (mode == QSLHIP_MODE_DENSE) {
            materialize(h);
            // block matrices from FP64 coefficients, rounded once to FP32
            std::vector<GateDesc> g64;
            g64.reserve(nops);
            for (uint64_t i = 0; i < nops; i++)
                g64.push_back(make_desc(8, ops[i].kind, ops[i].q0,
                                        is_gate2(ops[i].kind) ? ops[i].q1 : 0, ops[i].angle));
            std::vector<DensePass> dp;
            std::vector<DenseOp> dops;
            std::vector<float> mats;
            plan_dense_passes(h->n, g64, tuning().dense_f, dp, dops, mats);
            launch_dense_passes(h, tuning().dense_f, dp, dops, mats);
            return;
        }
        if (h->np && mode != QSLHIP_MODE_PER_GATE) {
            np_apply_gates(h, gates);
            return;
        }
        if (mode == QSLHIP_MODE_PER_GATE || h->np) {
            materialize(h);
            for (const GateDesc &g : gates) {
                if (h->np) np_launch_gate(h, g);
                else launch_gate(h, g);
            }
            return;
        }
        const Tuning &tu = tuning();
        int logt = tile_width(tu, h->n, h->prec);
        int low = h->prec == 8 ? tu.low_f64 : tu.low_f32;
        std::vector<TilePass> passes;
        std::vector<TileOp> tops;
        if (mode == QSLHIP_MODE_REORDER) {
            std::vector<uint32_t> order;
            plan_tile_passes_reorder(h->n, gates, logt, low, passes, tops, order);
        } else {
            plan_tile_passes(h->n, h->prec, gates, logt, low, passes, tops);
        }
        // a recorded basis state is synthesized by the first pass (or written
        // first); the last pass leaves the sampling sums of the result
        const bool sums = launch_tile_passes(h, passes, tops, nullptr,
                                             mode != QSLHIP_MODE_REORDER, h->pending_basis,
                                             sample_run_sums(h));
        h->pending_basis = qslhip_state::kNoBasis;
        h->sums_valid = sums;
        if (sums) {
            h->sums_tile = passes.back().tile_mask;
            h->sums_logt = passes.back().logt;
        }
    });
}

int qslhip_plan(unsigned nqubits, int prec, const qslhip_op *ops, uint64_t nops,
                uint64_t *npasses, uint64_t *tile_masks, uint64_t *op_counts,
                uint64_t cap)
{
    return qslhip_plan_mode(nqubits, prec, ops, nops, QSLHIP_MODE_FUSED, npasses, tile_masks,
                            op_counts, cap, nullptr);
}

int qslhip_plan_mode(unsigned nqubits, int prec, const qslhip_op *ops, uint64_t nops,
                     int mode, uint64_t *npasses, uint64_t *tile_masks,
                     uint64_t *op_counts, uint64_t cap, uint64_t *order)
{
    return guarded([&] {
        QH_REQUIRE(mode == QSLHIP_MODE_FUSED || mode == QSLHIP_MODE_REORDER, QSLHIP_EINVAL,
                   "plan mode must be FUSED or REORDER");
        QH_REQUIRE(npasses, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(ops || nops == 0, QSLHIP_EINVAL, "null op list");
        QH_REQUIRE(prec == 4 || prec == 8, QSLHIP_EINVAL, "prec must be 4 or 8");
        QH_REQUIRE(nqubits >= 1 && nqubits <= kMaxQubits, QSLHIP_EINVAL,
                   "nqubits must be in [1, 40]");
        std::vector<GateDesc> gates;
        for (uint64_t i = 0; i < nops; i++) {
            const qslhip_op &o = ops[i];
            bool ok1 = is_gate1(o.kind) && o.q0 < nqubits;
            bool ok2 = is_gate2(o.kind) && o.q0 < nqubits && o.q1 < nqubits && o.q0 != o.q1;
            QH_REQUIRE(ok1 || ok2, QSLHIP_EINVAL, "bad op " + std::to_string(i));
            gates.push_back(make_desc(prec, o.kind, o.q0, is_gate2(o.kind) ? o.q1 : 0,
                                      o.angle));
        }
        const Tuning &tu = tuning();
        int logt = tile_width(tu, nqubits, prec);
        int low = prec == 8 ? tu.low_f64 : tu.low_f32;
        std::vector<TilePass> passes;
        std::vector<TileOp> tops;
        std::vector<uint32_t> ord;
        if (mode == QSLHIP_MODE_REORDER) {
            plan_tile_passes_reorder(nqubits, gates, logt, low, passes, tops, ord);
        } else {
            plan_tile_passes(nqubits, prec, gates, logt, low, passes, tops);
            for (uint64_t i = 0; i < nops; i++) ord.push_back((uint32_t)i);
        }
        if (order)
            for (size_t i = 0; i < ord.size(); i++) order[i] = ord[i];
        *npasses = passes.size();
        for (uint64_t i = 0; i < passes.size() && i < cap; i++) {
            if (tile_masks) tile_masks[i] = passes[i].tile_mask;
            if (op_counts) op_counts[i] = passes[i].op_count;
        }
    });
}

int qslhip_jit_source(unsigned nqubits, int prec, const qslhip_op *ops, uint64_t nops, int mode,
                      uint64_t pass, char *buf, size_t cap, uint64_t *len)
{
    return guarded([&] {
        QH_REQUIRE(len, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(ops || nops == 0, QSLHIP_EINVAL, "null op list");
        QH_REQUIRE(prec == 4 || prec == 8, QSLHIP_EINVAL, "prec must be 4 or 8");
        QH_REQUIRE(mode == QSLHIP_MODE_FUSED || mode == QSLHIP_MODE_REORDER, QSLHIP_EINVAL,
                   "mode must be FUSED or REORDER");
        QH_REQUIRE(nqubits >= 1 && nqubits <= kMaxQubits, QSLHIP_EINVAL,
                   "nqubits must be in [1, 40]");
        std::vector<GateDesc> gates;
        for (uint64_t i = 0; i < nops; i++) {
            const qslhip_op &o = ops[i];
            bool ok1 = is_gate1(o.kind) && o.q0 < nqubits;
            bool ok2 = is_gate2(o.kind) && o.q0 < nqubits && o.q1 < nqubits && o.q0 != o.q1;
            QH_REQUIRE(ok1 || ok2, QSLHIP_EINVAL, "bad op " + std::to_string(i));
            gates.push_back(make_desc(prec, o.kind, o.q0, is_gate2(o.kind) ? o.q1 : 0,
                                      o.angle));
        }
        const Tuning &tu = tuning();
        const int logt = tile_width(tu, nqubits, prec);
        const int low = prec == 8 ? tu.low_f64 : tu.low_f32;
        const int threads = prec == 8 ? tu.rthreads_f64 : tu.rthreads_f32;
        std::vector<TilePass> passes;
        std::vector<TileOp> tops;
        std::vector<uint32_t> order;
        if (mode == QSLHIP_MODE_REORDER)
            plan_tile_passes_reorder(nqubits, gates, logt, low, passes, tops, order);
        else
            plan_tile_passes(nqubits, prec, gates, logt, low, passes, tops);
        QH_REQUIRE(pass < passes.size(), QSLHIP_EINVAL, "no such pass");
        const TilePass &p = passes[pass];
        QH_REQUIRE(reg_tile_ok(p.logt, threads, prec), QSLHIP_EINVAL,
                   "this pass runs on LDS tiles, not register tiles");
        std::vector<RegOp> rops;
        RegLayout lay;
        plan_reg_ops(tops.data() + p.op_first, p.op_count, p.logt, threads == 512 ? 9 : 8,
                     prec == 8 ? 16 : 8, rops, &lay);
        const std::string src = jit_pass_source(nqubits, prec, p, threads,
                                                mode != QSLHIP_MODE_REORDER, rops.data(),
                                                (uint32_t)rops.size(), lay);
        *len = src.size();
        if (buf && cap) {
            const size_t m = src.size() < cap - 1 ? src.size() : cap - 1;
            std::memcpy(buf, src.data(), m);
            buf[m] = 0;
        }
    });
}

int qslhip_jit_check(const char *src, const char *arch, char *log, size_t cap)
{
    return guarded([&] {
        QH_REQUIRE(src && arch, QSLHIP_EINVAL, "null argument");
        std::string lg;
        const bool ok = jit_check_source(src, arch, lg);
        if (log && cap) {
            const size_t m = lg.size() < cap - 1 ? lg.size() : cap - 1;
            std::memcpy(log, lg.data(), m);
            log[m] = 0;
        }
        QH_REQUIRE(ok, QSLHIP_EINVAL, "run-time compile failed: " + lg.substr(0, 400));
    });
}

int qslhip_plan_dense(unsigned nqubits, const qslhip_op *ops, uint64_t nops, int F,
                      uint64_t *npasses, uint64_t *nops_out, uint64_t *nfloats,
                      uint64_t *pass_info, int64_t *op_info, float *mats,
                      uint64_t cap_passes, uint64_t cap_ops, uint64_t cap_floats)
{
    return guarded([&] {
        QH_REQUIRE(npasses && nops_out && nfloats, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(ops || nops == 0, QSLHIP_EINVAL, "null op list");
        QH_REQUIRE(F == 4 || F == 5, QSLHIP_EINVAL, "fusion width must be 4 or 5");
        QH_REQUIRE(nqubits >= (unsigned)kDenseTile && nqubits <= kMaxQubits, QSLHIP_EINVAL,
                   "dense plans need 12..40 qubits");
        std::vector<GateDesc> g64;
        for (uint64_t i = 0; i < nops; i++) {
            const qslhip_op &o = ops[i];
            bool ok1 = is_gate1(o.kind) && o.q0 < nqubits;
            bool ok2 = is_gate2(o.kind) && o.q0 < nqubits && o.q1 < nqubits && o.q0 != o.q1;
            QH_REQUIRE(ok1 || ok2, QSLHIP_EINVAL, "bad op " + std::to_string(i));
            g64.push_back(make_desc(8, o.kind, o.q0, is_gate2(o.kind) ? o.q1 : 0, o.angle));
        }
        std::vector<DensePass> dp;
        std::vector<DenseOp> dops;
        std::vector<float> m;
        plan_dense_passes(nqubits, g64, F, dp, dops, m);
        *npasses = dp.size();
        *nops_out = dops.size();
        *nfloats = m.size();
        for (uint64_t i = 0; pass_info && i < dp.size() && i < cap_passes; i++) {
            pass_info[3 * i] = dp[i].tile_mask;
            pass_info[3 * i + 1] = dp[i].op_first;
            pass_info[3 * i + 2] = dp[i].op_count;
        }
        for (uint64_t i = 0; op_info && i < dops.size() && i < cap_ops; i++) {
            const DenseOp &o = dops[i];
            int64_t *r = op_info + 9 * i;
            r[0] = o.kind;
            r[1] = o.gidx;
            r[2] = o.kind ? o.nq : 0;
            r[3] = o.kind ? (int64_t)o.mat * 2 : -1;
            // tile positions -> qubits through the pass's tile mask
            uint64_t mask = 0;
            for (const DensePass &P : dp)
                if (i >= P.op_first && i < P.op_first + P.op_count) mask = P.tile_mask;
            for (int j = 0; j < 5; j++) {
                int64_t q = -1;
                if (o.kind && j < o.nq) {
                    int pos = o.pos[j], seen = 0;
                    for (unsigned b = 0; b < 64; b++)
                        if ((mask >> b) & 1ull) {
                            if (seen == pos) { q = b; break; }
                            seen++;
                        }
                }
                r[4 + j] = q;
            }
        }
        if (mats && cap_floats) std::memcpy(mats, m.data(), std::min<uint64_t>(cap_floats, m.size()) * sizeof(float));
    });
}

int qslhip_prob(qslhip_t h, unsigned targ, unsigned outcome, double *p)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(p, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(targ < h->n && outcome <= 1, QSLHIP_EINVAL, "bad qubit or outcome");
        *p = h->np ? np_prob(h, targ, outcome) : device_prob(h, targ, outcome);
    });
}

int qslhip_collapse(qslhip_t h, unsigned targ, unsigned outcome, double factor)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(targ < h->n && outcome <= 1, QSLHIP_EINVAL, "bad qubit or outcome");
        if (h->np) np_collapse(h, targ, outcome, factor);
        else launch_collapse(h, targ, outcome, factor);
    });
}

// u is compared in the handle's Fp, as the reference's `Fp rand`.
static double as_fp(qslhip_t h, double u)
{
    return h->prec == 8 ? u : (double)(float)u;
}

int qslhip_measure(qslhip_t h, unsigned targ, double u, int *outcome)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(outcome, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(targ < h->n, QSLHIP_EINVAL, "qubit index out of range");
        // measure.cpp:81-107, with prob0 and the factor in FP64
        double prob0 = h->np ? np_prob(h, targ, 0) : device_prob(h, targ, 0);
        double uu = as_fp(h, u);
        int o;
        double factor;
        if (uu < prob0) {
            o = 0;
            factor = 1.0 / std::sqrt(prob0);
        } else {
            o = 1;
            factor = 1.0 / std::sqrt(1.0 - prob0);
        }
        if (h->np) np_collapse(h, targ, (unsigned)o, factor);
        else launch_collapse(h, targ, (unsigned)o, factor);
        QH_CHECK(hipStreamSynchronize(h->stream));
        *outcome = o;
    });
}

int qslhip_postselect(qslhip_t h, unsigned targ, unsigned outcome, double *p)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(p, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(targ < h->n && outcome <= 1, QSLHIP_EINVAL, "bad qubit or outcome");
        // measure.cpp:149-161
        double pr = h->np ? np_prob(h, targ, outcome) : device_prob(h, targ, outcome);
        double factor = 1.0 / std::sqrt(pr);
        if (h->np) np_collapse(h, targ, outcome, factor);
        else launch_collapse(h, targ, outcome, factor);
        QH_CHECK(hipStreamSynchronize(h->stream));
        *p = pr;
    });
}

int qslhip_sample_all(qslhip_t h, const double *u, uint64_t nshots, uint64_t *out)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE((u && out) || nshots == 0, QSLHIP_EINVAL, "null array");
        std::vector<double> uf;
        const double *src = u;
        if (h->prec == 4 && nshots) {
            uf.assign(u, u + nshots);
            for (double &x : uf) x = (double)(float)x;
            src = uf.data();
        }
        if (h->np) np_sample(h, src, nshots, out);
        else device_sample(h, src, nshots, out);
    });
}

// ---- resizeable handles (src/qubits/resize/) ------------------------------------

static void require_resizable(qslhip_t h)
{
    QH_REQUIRE(h->resizable, QSLHIP_EKIND,
               "qubits are added and removed only on a resizeable handle (qslhip_create_resize)");
}

int qslhip_capacity(qslhip_t h, uint64_t *cap)
{
    return guarded([&] {
        require_handle(h);
        QH_REQUIRE(cap, QSLHIP_EINVAL, "null output");
        *cap = h->cap;
    });
}

int qslhip_resize_set_state(qslhip_t h, const void *host, uint64_t count)
{
    return guarded([&] {
        require_handle(h);
        require_resizable(h);
        QH_REQUIRE(host, QSLHIP_EINVAL, "null host pointer");
        // resize.cpp:70-81 (checkStateSize, quantum.cpp:108-131)
        QH_REQUIRE(count >= 2 && (count & (count - 1)) == 0, QSLHIP_ESIZE,
                   "State vector size is not a power of two");
        QH_REQUIRE(count <= (1ull << kMaxQubits), QSLHIP_EINVAL, "state too large");
        if (count > h->cap) {
            void *p = nullptr;
            hipError_t e = hipMalloc(&p, count * h->amp_bytes());
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(QSLHIP_ENOMEM, "hipMalloc of " +
                                               std::to_string(count * h->amp_bytes()) +
                                               " bytes for the state failed");
            }
            QH_CHECK(hipStreamSynchronize(h->stream));
            QH_CHECK(hipFree(h->amps));
            h->amps = p;
            h->cap = count;
        }
        h->n = (unsigned)__builtin_ctzll(count);
        h->namps = count;
        h->pending_basis = qslhip_state::kNoBasis;
        h->sums_valid = false;
        QH_CHECK(hipMemcpyAsync(h->amps, host, count * h->amp_bytes(), hipMemcpyHostToDevice,
                                h->stream));
        QH_CHECK(hipStreamSynchronize(h->stream));
    });
}

int qslhip_add_qubit(qslhip_t h, unsigned targ)
{
    return guarded([&] {
        require_handle(h);
        require_resizable(h);
        // resize.cpp:107-147; targ == n appends (appendQubit, :101-105)
        QH_REQUIRE(targ <= h->n, QSLHIP_EINVAL, "qubit index out of range");
        QH_REQUIRE(h->n + 1 <= kMaxQubits, QSLHIP_EINVAL, "too many qubits");
        require_state(h);
        resize_add_qubit(h, targ);
        QH_CHECK(hipStreamSynchronize(h->stream));
    });
}

int qslhip_collapse_out(qslhip_t h, unsigned targ, unsigned outcome, double factor)
{
    return guarded([&] {
        require_handle(h);
        require_resizable(h);
        QH_REQUIRE(targ < h->n && outcome <= 1, QSLHIP_EINVAL, "bad qubit or outcome");
        QH_REQUIRE(h->n >= 2, QSLHIP_EINVAL, "cannot remove the last qubit");
        require_state(h);
        // measure.cpp:62-94
        resize_collapse_out(h, targ, outcome, factor);
        QH_CHECK(hipStreamSynchronize(h->stream));
    });
}

int qslhip_measure_out(qslhip_t h, unsigned targ, double u, int *outcome)
{
    return guarded([&] {
        require_handle(h);
        require_resizable(h);
        QH_REQUIRE(outcome, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(targ < h->n, QSLHIP_EINVAL, "qubit index out of range");
        QH_REQUIRE(h->n >= 2, QSLHIP_EINVAL, "cannot remove the last qubit");
        require_state(h);
        // measure.cpp:157-190, with prob0 and the factor in FP64 as measure
        double prob0 = device_prob(h, targ, 0);
        double uu = as_fp(h, u);
        int o;
        double factor;
        if (uu < prob0) {
            o = 0;
            factor = 1.0 / std::sqrt(prob0);
        } else {
            o = 1;
            factor = 1.0 / std::sqrt(1.0 - prob0);
        }
        resize_collapse_out(h, targ, (unsigned)o, factor);
        QH_CHECK(hipStreamSynchronize(h->stream));
        *outcome = o;
    });
}

int qslhip_postselect_out(qslhip_t h, unsigned targ, unsigned outcome, double *p)
{
    return guarded([&] {
        require_handle(h);
        require_resizable(h);
        QH_REQUIRE(p, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(targ < h->n && outcome <= 1, QSLHIP_EINVAL, "bad qubit or outcome");
        QH_REQUIRE(h->n >= 2, QSLHIP_EINVAL, "cannot remove the last qubit");
        require_state(h);
        // measure.cpp:235-250
        double pr = device_prob(h, targ, outcome);
        resize_collapse_out(h, targ, outcome, 1.0 / std::sqrt(pr));
        QH_CHECK(hipStreamSynchronize(h->stream));
        *p = pr;
    });
}

int qslhip_reallocate(qslhip_t h)
{
    return guarded([&] {
        require_handle(h);
        require_resizable(h);
        // resize.cpp:93-98: trim the allocation to the state
        if (h->cap == h->namps) return;
        require_state_ro(h);
        void *p = nullptr;
        QH_CHECK(hipMalloc(&p, h->namps * h->amp_bytes()));
        QH_CHECK(hipMemcpyAsync(p, h->amps, h->namps * h->amp_bytes(), hipMemcpyDeviceToDevice,
                                h->stream));
        QH_CHECK(hipStreamSynchronize(h->stream));
        QH_CHECK(hipFree(h->amps));
        h->amps = p;
        h->cap = h->namps;
    });
}

int qslhip_jit_wait(void)
{
    return guarded([&] { jit_wait_idle(); });
}

int qslhip_sample_all_noalloc(qslhip_t h, const double *u, uint64_t nshots, uint64_t *out)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE((u && out) || nshots == 0, QSLHIP_EINVAL, "null array");
        device_sample_noalloc(h, u, nshots, out);
    });
}

int qslhip_sample_all2(qslhip_t h, const double *u, uint64_t nshots,
                       uint64_t *counts, uint64_t ncounts)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(counts && (u || nshots == 0), QSLHIP_EINVAL, "null array");
        QH_REQUIRE(ncounts == (1ull << h->n), QSLHIP_ESIZE, "ncounts must be 2^nqubits");
        std::vector<double> uf(u, u + nshots);
        if (h->prec == 4)
            for (double &x : uf) x = (double)(float)x;
        std::vector<uint64_t> out(nshots);
        if (h->np) np_sample(h, uf.data(), nshots, out.data(), true);
        else device_sample(h, uf.data(), nshots, out.data(), true);
        std::memset(counts, 0, ncounts * sizeof(uint64_t));
        for (uint64_t x : out)
            if (x != ~0ull) counts[x]++;
    });
}

int qslhip_measure_all(qslhip_t h, double u, uint64_t *outcome)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(outcome, QSLHIP_EINVAL, "null output");
        double uu = as_fp(h, u);
        uint64_t o = 0;
        if (h->np) np_sample(h, &uu, 1, &o);
        else device_sample(h, &uu, 1, &o);
        // measure.cpp:110-123: collapse onto the outcome
        if (h->np) launch_fill_basis(h, np_rank_host(o));
        else h->pending_basis = o;  // the collapsed state, written when next read
        h->sums_valid = false;
        QH_CHECK(hipStreamSynchronize(h->stream));
        *outcome = o;
    });
}

int qslhip_sample(qslhip_t h, unsigned targ, const double *u, uint64_t nshots,
                  uint64_t *counts)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(counts && (u || nshots == 0), QSLHIP_EINVAL, "null array");
        QH_REQUIRE(targ < h->n, QSLHIP_EINVAL, "qubit index out of range");
        // measure.cpp:191-216: one prob, then u < prob0 counts as 0
        double prob0 = h->np ? np_prob(h, targ, 0) : device_prob(h, targ, 0);
        uint64_t c0 = 0;
        for (uint64_t i = 0; i < nshots; i++)
            if (as_fp(h, u[i]) < prob0) c0++;
        counts[0] = c0;
        counts[1] = nshots - c0;
    });
}

int qslhip_norm(qslhip_t h, double *norm)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(norm, QSLHIP_EINVAL, "null output");
        double s = h->np ? np_prob(h, 0, 0) + np_prob(h, 0, 1) : device_norm2(h);
        *norm = std::sqrt(s);
    });
}

static void validate_terms(qslhip_t h, const qslhip_pauli *terms, uint64_t nterms)
{
    QH_REQUIRE(terms || nterms == 0, QSLHIP_EINVAL, "null terms");
    const uint64_t lim = h->n >= 64 ? ~0ull : (1ull << h->n);
    for (uint64_t t = 0; t < nterms; t++)
        QH_REQUIRE(terms[t].xmask < lim && terms[t].zmask < lim, QSLHIP_EINVAL,
                   "Pauli term on a qubit outside the register");
}

int qslhip_expval(qslhip_t h, const qslhip_pauli *terms, uint64_t nterms, double *value)
{
    return guarded([&] {
        require_state_ro(h);
        QH_REQUIRE(value, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(!h->np, QSLHIP_EKIND, "expectation values need a full-vector handle");
        validate_terms(h, terms, nterms);
        *value = expval(h, terms, nterms);
    });
}

int qslhip_expval_grad(qslhip_t h, const qslhip_op *ops, uint64_t nops,
                       const qslhip_pauli *terms, uint64_t nterms, double *energy,
                       double *grads)
{
    return guarded([&] {
        require_state(h);
        QH_REQUIRE(energy, QSLHIP_EINVAL, "null output");
        QH_REQUIRE(ops || nops == 0, QSLHIP_EINVAL, "null op list");
        QH_REQUIRE(!h->np, QSLHIP_EKIND, "gradients need a full-vector handle");
        validate_terms(h, terms, nterms);
        std::vector<GateDesc> gates;
        gates.reserve(nops);
        for (uint64_t i = 0; i < nops; i++) {
            const qslhip_op &o = ops[i];
            validate_gate(h, o.kind, o.q0, o.q1);
            gates.push_back(make_desc(h->prec, o.kind, o.q0,
                                      is_gate2(o.kind) ? o.q1 : 0, o.angle));
        }
        adjoint_plain(h, gates, terms, nterms, energy, grads);
    });
}

int qslhip_uniforms(uint64_t seed, uint64_t count, double *out)
{
    return guarded([&] {
        QH_REQUIRE(out || count == 0, QSLHIP_EINVAL, "null output");
        // splitmix64 of the seed (Steele, Lea & Flood 2014)
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        std::mt19937_64 gen(z);
        for (uint64_t i = 0; i < count; i++)
            out[i] = (double)(gen() >> 11) * 0x1.0p-53;
    });
}

}  // extern "C"
