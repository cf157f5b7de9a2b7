// Out-of-core contribution store: the reference's ContributionStore
// (src/amls/contribution_store.hpp:23-124), for update blocks that do not
// stay resident.  Each target (an ancestor block id) has one append-only
// file <dir>/t<target>.bin of records
//     int64 rows, int64 cols, rows*cols doubles (column-major),
// native byte order, the layout the reference writes (:34-49), so stores
// written by either side read back on the other.
//
// What differs from the reference is where the blocks live.  With a device
// (device >= 0) the blocks are HBM pointers: append streams a block to the
// file through two pinned host buffers (the device-to-host copy of one piece
// overlaps the file write of the previous one), and read_sum streams the
// records back and sums them on the GPU into the caller's device block, one
// IEEE addition per element and record in append order, which is the
// reference's out = first; out += next (:62-84) bit for bit.  Without a
// device (device < 0) the pointers are host memory and the same code paths
// run with host copies.
//
// Semantics kept from the reference: appends to one target are serialised
// (striped locks, :36-37) and may come from several threads; read_sum of a
// missing file reports "not found" and leaves the output untouched (:57-58);
// a truncated trailing record is ignored (:67-77); consume deletes the file
// (:79-85); reset deletes every file this store wrote (:96-103); disk_bytes
// counts the payload bytes appended since the last reset (:48, :106).  Added:
// a record whose shape differs from the output block is an error (Eigen
// would assert), and shape() reads a target's record headers.
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "amls_hip.h"
#include "kernels/launch.hpp"

namespace {

struct StoreError : std::runtime_error {
    int code;
    StoreError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define STORE_HIP(expr)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw StoreError(e_ == hipErrorOutOfMemory ? AMLS_E_OOM : AMLS_E_HIP,               \
                             std::string("HIP error '") + hipGetErrorString(e_) + "' at " #expr); \
    } while (0)

// A contiguous piece of a rows x cols column-major block that fits `cap`
// elements: whole columns [c0, c0 + nc) when a column fits, else rows
// [r0, r0 + nr) of one column.  Pieces in order cover the block in its
// column-major order, so each is one contiguous range of the record.
struct Piece {
    int64_t c0, nc, r0, nr;
};

std::vector<Piece> piecesOf(int64_t rows, int64_t cols, int64_t cap) {
    std::vector<Piece> out;
    if (rows <= 0 || cols <= 0) return out;
    if (rows <= cap) {
        const int64_t g = std::max<int64_t>(1, cap / rows);
        for (int64_t c = 0; c < cols; c += g) out.push_back({c, std::min(g, cols - c), 0, rows});
    } else {
        for (int64_t c = 0; c < cols; ++c)
            for (int64_t r = 0; r < rows; r += cap) out.push_back({c, 1, r, std::min(cap, rows - r)});
    }
    return out;
}

struct File {
    std::FILE* f = nullptr;
    File(const std::string& p, const char* mode) : f(std::fopen(p.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool close() {
        const bool ok = f && std::fclose(f) == 0;
        f = nullptr;
        return ok;
    }
};

}  // namespace

struct amls_store {
    std::string dir;
    int device = -1;
    int64_t cap = 0;                        // elements per staging buffer
    std::array<std::mutex, 64> stripes;     // appends / reads of one target serialise
    std::mutex setMu;
    std::unordered_set<int64_t> targets;
    std::atomic<int64_t> bytes{0};
    // one staging set per store: device transfers serialise on devMu
    std::mutex devMu;
    hipStream_t stream = nullptr;
    double* pin[2] = {nullptr, nullptr};
    double* dbuf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<double> hbuf;               // host-mode read buffer
    std::mutex errMu;
    std::string err;

    std::string path(int64_t t) const { return dir + "/t" + std::to_string(t) + ".bin"; }
    std::mutex& stripe(int64_t t) { return stripes[static_cast<uint64_t>(t) % stripes.size()]; }

    // staging on first device use, with the store's device current
    void ensureDevice() {
        STORE_HIP(hipSetDevice(device));
        if (stream) return;
        // a blocking stream: ordered after work the caller put on the null
        // stream (work on other streams the caller must finish first)
        STORE_HIP(hipStreamCreate(&stream));
        for (int b = 0; b < 2; ++b) {
            STORE_HIP(hipHostMalloc(reinterpret_cast<void**>(&pin[b]), cap * sizeof(double),
                                    hipHostMallocDefault));
            STORE_HIP(hipMalloc(reinterpret_cast<void**>(&dbuf[b]), cap * sizeof(double)));
            STORE_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
    }
    void releaseDevice() {
        if (!stream) return;
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
        for (int b = 0; b < 2; ++b) {
            if (pin[b]) (void)hipHostFree(pin[b]);
            if (dbuf[b]) (void)hipFree(dbuf[b]);
            if (ev[b]) (void)hipEventDestroy(ev[b]);
        }
        (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    void append(int64_t target, const double* m, int64_t rows, int64_t cols, int64_t ld) {
        if (rows < 0 || cols < 0) throw StoreError(AMLS_E_ARG, "negative block size");
        if (rows * cols > 0 && !m) throw StoreError(AMLS_E_ARG, "null block");
        if (cols > 1 && ld < rows) throw StoreError(AMLS_E_ARG, "ld < rows");
        if (ld < rows) ld = rows;
        const std::string p = path(target);
        std::lock_guard<std::mutex> lk(stripe(target));
        File f(p, "ab");
        if (!f.f) throw StoreError(AMLS_E_ARG, "cannot open " + p + " for appending");
        const int64_t hdr[2] = {rows, cols};
        if (std::fwrite(hdr, sizeof(int64_t), 2, f.f) != 2) throw StoreError(AMLS_E_INTERNAL, "write failed: " + p);
        const auto pcs = piecesOf(rows, cols, cap);
        auto put = [&](const double* src, int64_t count) {
            if (count > 0 && std::fwrite(src, sizeof(double), count, f.f) != static_cast<size_t>(count))
                throw StoreError(AMLS_E_INTERNAL, "write failed: " + p);
        };
        if (device < 0) {
            for (const Piece& q : pcs)
                for (int64_t c = q.c0; c < q.c0 + q.nc; ++c) put(m + q.r0 + c * ld, q.nr);
        } else {
            std::lock_guard<std::mutex> dk(devMu);
            ensureDevice();
            // piece i+1 copies while piece i is written; a buffer is refilled
            // only after its previous piece has been written
            auto fetch = [&](size_t i) {
                const Piece& q = pcs[i];
                const int b = static_cast<int>(i & 1);
                STORE_HIP(hipMemcpy2DAsync(pin[b], q.nr * sizeof(double), m + q.r0 + q.c0 * ld,
                                           ld * sizeof(double), q.nr * sizeof(double), q.nc,
                                           hipMemcpyDeviceToHost, stream));
                STORE_HIP(hipEventRecord(ev[b], stream));
            };
            if (!pcs.empty()) fetch(0);
            for (size_t i = 0; i < pcs.size(); ++i) {
                if (i + 1 < pcs.size()) fetch(i + 1);
                const int b = static_cast<int>(i & 1);
                STORE_HIP(hipEventSynchronize(ev[b]));
                put(pin[b], pcs[i].nr * pcs[i].nc);
            }
        }
        if (!f.close()) throw StoreError(AMLS_E_INTERNAL, "close failed: " + p);
        {
            std::lock_guard<std::mutex> sk(setMu);
            targets.insert(target);
        }
        bytes += rows * cols * static_cast<int64_t>(sizeof(double));
    }

    // the records of a target: (offset of the data, rows, cols) of each
    // complete record, in append order; false when there is no file
    bool scan(const std::string& p, std::vector<std::array<int64_t, 3>>& recs) {
        std::error_code ec;
        const auto size = std::filesystem::file_size(p, ec);
        File f(p, "rb");
        if (!f.f || ec) return false;
        const int64_t total = static_cast<int64_t>(size);
        int64_t off = 0;
        while (off + 16 <= total) {
            int64_t hdr[2];
            if (std::fseek(f.f, off, SEEK_SET) != 0 || std::fread(hdr, sizeof(int64_t), 2, f.f) != 2) break;
            if (hdr[0] < 0 || hdr[1] < 0) break;
            const int64_t need = hdr[0] * hdr[1] * static_cast<int64_t>(sizeof(double));
            if (off + 16 + need > total) break;   // truncated record: ignored, as in the reference
            recs.push_back({off + 16, hdr[0], hdr[1]});
            off += 16 + need;
        }
        return true;
    }

    bool readSum(int64_t target, double* out, int64_t rows, int64_t cols, int64_t ld, bool consume) {
        if (rows < 0 || cols < 0) throw StoreError(AMLS_E_ARG, "negative block size");
        if (cols > 1 && ld < rows) throw StoreError(AMLS_E_ARG, "ld < rows");
        if (ld < rows) ld = rows;
        const std::string p = path(target);
        std::lock_guard<std::mutex> lk(stripe(target));
        std::vector<std::array<int64_t, 3>> recs;
        if (!scan(p, recs)) return false;
        for (const auto& r : recs)
            if (r[1] != rows || r[2] != cols)
                throw StoreError(AMLS_E_ARG, "record of target " + std::to_string(target) + " is " +
                                                 std::to_string(r[1]) + " x " + std::to_string(r[2]) +
                                                 ", the output block " + std::to_string(rows) + " x " +
                                                 std::to_string(cols));
        if (!recs.empty() && rows * cols > 0 && !out) throw StoreError(AMLS_E_ARG, "null output block");
        File f(p, "rb");
        if (!f.f) throw StoreError(AMLS_E_INTERNAL, "cannot reopen " + p);
        const auto pcs = piecesOf(rows, cols, cap);
        auto get = [&](double* dst, int64_t off, int64_t count) {
            if (std::fseek(f.f, off, SEEK_SET) != 0 ||
                std::fread(dst, sizeof(double), count, f.f) != static_cast<size_t>(count))
                throw StoreError(AMLS_E_INTERNAL, "read failed: " + p);
        };
        // element offset of a piece inside its record
        auto at = [&](const Piece& q) { return (q.c0 * rows + q.r0) * static_cast<int64_t>(sizeof(double)); };
        if (device < 0) {
            hbuf.resize(static_cast<size_t>(cap));
            for (size_t k = 0; k < recs.size(); ++k)
                for (const Piece& q : pcs) {
                    get(hbuf.data(), recs[k][0] + at(q), q.nr * q.nc);
                    for (int64_t c = 0; c < q.nc; ++c) {
                        double* o = out + q.r0 + (q.c0 + c) * ld;
                        const double* s = hbuf.data() + c * q.nr;
                        if (k == 0)
                            std::memcpy(o, s, q.nr * sizeof(double));
                        else
                            for (int64_t r = 0; r < q.nr; ++r) o[r] += s[r];
                    }
                }
        } else {
            std::lock_guard<std::mutex> dk(devMu);
            ensureDevice();
            size_t i = 0;   // running piece count: buffer i & 1
            for (size_t k = 0; k < recs.size(); ++k)
                for (const Piece& q : pcs) {
                    const int b = static_cast<int>(i & 1);
                    // the buffer's previous copy (and the addition reading its
                    // device twin) are done before it is refilled
                    if (i >= 2) STORE_HIP(hipEventSynchronize(ev[b]));
                    get(pin[b], recs[k][0] + at(q), q.nr * q.nc);
                    double* o = out + q.r0 + q.c0 * ld;
                    if (k == 0) {
                        STORE_HIP(hipMemcpy2DAsync(o, ld * sizeof(double), pin[b], q.nr * sizeof(double),
                                                   q.nr * sizeof(double), q.nc, hipMemcpyHostToDevice,
                                                   stream));
                    } else {
                        STORE_HIP(hipMemcpyAsync(dbuf[b], pin[b], q.nr * q.nc * sizeof(double),
                                                 hipMemcpyHostToDevice, stream));
                        amls::launch_add2d(o, ld, dbuf[b], q.nr, q.nr, q.nc, stream);
                        STORE_HIP(hipGetLastError());
                    }
                    STORE_HIP(hipEventRecord(ev[b], stream));
                    ++i;
                }
            STORE_HIP(hipStreamSynchronize(stream));
        }
        f.close();
        if (consume) {
            std::error_code ec;
            std::filesystem::remove(p, ec);
            std::lock_guard<std::mutex> sk(setMu);
            targets.erase(target);
        }
        return true;
    }

    void remove(int64_t target) {
        std::lock_guard<std::mutex> lk(stripe(target));
        std::error_code ec;
        std::filesystem::remove(path(target), ec);
        std::lock_guard<std::mutex> sk(setMu);
        targets.erase(target);
    }

    void reset() {
        std::lock_guard<std::mutex> sk(setMu);
        for (int64_t t : targets) {
            std::error_code ec;
            std::filesystem::remove(path(t), ec);
        }
        targets.clear();
        bytes = 0;
    }
};

namespace {

template <class F>
amls_status storeGuarded(amls_store* st, F&& f) {
    if (!st) return AMLS_E_ARG;
    try {
        f(*st);
        return AMLS_OK;
    } catch (const StoreError& e) {
        std::lock_guard<std::mutex> lk(st->errMu);
        st->err = e.what();
        return static_cast<amls_status>(e.code);
    } catch (const std::bad_alloc&) {
        std::lock_guard<std::mutex> lk(st->errMu);
        st->err = "host allocation failed";
        return AMLS_E_OOM;
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(st->errMu);
        st->err = e.what();
        return AMLS_E_INTERNAL;
    }
}

}  // namespace

extern "C" {

amls_status amls_hip_store_open(const char* dir, int32_t device, int64_t staging_bytes,
                                amls_store** out) {
    if (!dir || !*dir || !out) return AMLS_E_ARG;
    *out = nullptr;
    if (staging_bytes < 0) return AMLS_E_ARG;
    try {
        std::filesystem::create_directories(dir);
        if (device >= 0) {
            int count = 0;
            if (hipGetDeviceCount(&count) != hipSuccess || device >= count) return AMLS_E_STATE;
        }
        auto* s = new amls_store();
        s->dir = dir;
        s->device = device < 0 ? -1 : device;
        const int64_t sb = staging_bytes > 0 ? staging_bytes : (int64_t(64) << 20);
        s->cap = std::max<int64_t>(1, sb / static_cast<int64_t>(sizeof(double)));
        *out = s;
        return AMLS_OK;
    } catch (const std::exception&) {
        return AMLS_E_ARG;
    }
}

amls_status amls_hip_store_close(amls_store* st) {
    if (!st) return AMLS_E_ARG;
    st->reset();
    st->releaseDevice();
    delete st;
    return AMLS_OK;
}

const char* amls_hip_store_last_error(amls_store* st) {
    if (!st) return "null store";
    std::lock_guard<std::mutex> lk(st->errMu);
    return st->err.c_str();
}

amls_status amls_hip_store_append(amls_store* st, int64_t target, const double* m, int64_t rows,
                                  int64_t cols, int64_t ld) {
    return storeGuarded(st, [&](amls_store& s) { s.append(target, m, rows, cols, ld); });
}

amls_status amls_hip_store_read_sum(amls_store* st, int64_t target, double* out, int64_t rows,
                                    int64_t cols, int64_t ld, int32_t consume, int32_t* found) {
    return storeGuarded(st, [&](amls_store& s) {
        const bool ok = s.readSum(target, out, rows, cols, ld, consume != 0);
        if (found) *found = ok ? 1 : 0;
    });
}

amls_status amls_hip_store_shape(amls_store* st, int64_t target, int64_t* rows, int64_t* cols,
                                 int64_t* records) {
    return storeGuarded(st, [&](amls_store& s) {
        std::lock_guard<std::mutex> lk(s.stripe(target));
        std::vector<std::array<int64_t, 3>> recs;
        const bool ok = s.scan(s.path(target), recs);
        if (rows) *rows = recs.empty() ? 0 : recs[0][1];
        if (cols) *cols = recs.empty() ? 0 : recs[0][2];
        if (records) *records = ok ? static_cast<int64_t>(recs.size()) : -1;
    });
}

amls_status amls_hip_store_remove(amls_store* st, int64_t target) {
    return storeGuarded(st, [&](amls_store& s) { s.remove(target); });
}

amls_status amls_hip_store_reset(amls_store* st) {
    return storeGuarded(st, [&](amls_store& s) { s.reset(); });
}

amls_status amls_hip_store_disk_bytes(amls_store* st, double* bytes) {
    return storeGuarded(st, [&](amls_store& s) {
        if (!bytes) throw StoreError(AMLS_E_ARG, "null output");
        *bytes = static_cast<double>(s.bytes.load());
    });
}

}  // extern "C"
