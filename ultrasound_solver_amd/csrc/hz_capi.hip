// C ABI of the Helmholtz library: operator handle, stencil applies, the two
// GMRES forms (plain FP64 with the reference's restart semantics, and
// complex64 inner solves under FP64 iterative refinement), the source
// columns and trilinear sampling.  Declarations: include/hz.h.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#include <thread>
#include <mutex>
#include <map>
#include <atomic>

#include "../../include/hz.h"
#include "hz_basis.hpp"
#include "hz_grad.hpp"
#include "hz_krylov.hpp"
#include "hz_mg.hpp"
#include "hz_slab.hpp"
#include "hz_materials.hpp"
#include "hz_source.hpp"
#include "hz_stencil.hpp"
#include "hz_surface.hpp"

using namespace hz;

namespace {

thread_local std::string g_error;

// HZ_VERBOSE=1 prints the residual at every restart-cycle top to stderr.
bool verbose() {
    static const bool on = [] {
        const char *v = std::getenv("HZ_VERBOSE");
        return v && v[0] && v[0] != '0';
    }();
    return on;
}

// Threshold of the two-step blocks' predicted intra-block coefficient (see
// k_block_alpha); HZ_BLOCK_GUARD overrides it per solve (1 sends every
// block through the measured fix-up, for tests).
double block_guard() {
    const char *v = std::getenv("HZ_BLOCK_GUARD");
    if (!v || !v[0]) return BLOCK_GUARD;
    const double g = std::atof(v);
    return g > 0 ? g : BLOCK_GUARD;
}

struct HipFailure : std::runtime_error {
    hipError_t code;
    HipFailure(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};
struct BadArgument : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HZ_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw HipFailure(e_, std::string(#call) + ": " + hipGetErrorString(e_) + " (line " +      \
                                     std::to_string(__LINE__) + ")");                                  \
    } while (0)

#define HZ_LAUNCHED() HZ_HIP(hipGetLastError())

// Owning device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t nbytes) {
        if (nbytes <= bytes) return;
        release();
        HZ_HIP(hipMalloc(&p, std::max<size_t>(nbytes, 256)));
        bytes = nbytes;
    }
    template <typename U>
    U *as() const { return static_cast<U *>(p); }
};

template <typename U>
void upload(DevBuf &dst, const U *src, size_t count) {
    dst.ensure(count * sizeof(U));
    HZ_HIP(hipMemcpy(dst.p, src, count * sizeof(U), hipMemcpyHostToDevice));
}

int64_t blocks_for(int64_t n, int64_t per_block, int64_t cap) {
    return std::max<int64_t>(1, std::min<int64_t>(cap, (n + per_block - 1) / per_block));
}

// One multigrid level (complex64 only: the preconditioner runs in the mixed form).
struct MgLevel {
    int nx = 0, ny = 0, nz = 0;
    int64_t n = 0;
    DevBuf rho, kap, a[3], q[3], face[6], dinv;
    OpView<float> v{};
    TileMap tm{};
    // Slab tile shapes of this level's smoother, residual and first-sweep
    // residual (the last holds more live registers: two planes per thread).
    int shape[3] = {0, 0, 1};
    DevBuf bb, xx, tt;  // k-slot work vectors: right-hand side, iterate, scratch
};

struct Multigrid {
    std::vector<std::unique_ptr<MgLevel>> levels;
    int pre = 1, post = 1, coarse_sweeps = 30;
    int post_coarse = 1;  // post-sweeps on the levels between the finest and the coarsest
    int cycle = 1;  // coarse-grid visits per level below the finest (1: V, 2: W)
    float weight = 0.7f;
    double shift = 0.5;
    int kcap = 0;
};

// Per-scene argument tables of fused launches (see Fuser): written on the
// host by a recording pass over a launch sequence, then uploaded once and
// read by every launch (or graph replay) of that sequence.
struct FusePlan {
    std::vector<char> host;
    std::vector<size_t> offs;  // table of each fused launch, in launch order
    DevBuf dev;
    size_t next = 0;
    bool recording = true;
    // On the solve's stream (no legacy-stream copy, which would order itself
    // after every blocking stream, including one another thread captures).
    void upload(hipStream_t s) {
        if (!host.empty()) {
            dev.ensure(host.size());
            HZ_HIP(hipMemcpyAsync(dev.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
            HZ_HIP(hipStreamSynchronize(s));
        }
        recording = false;
        next = 0;
    }
};

// Launch sequences kept for replay, keyed by everything their launches bake
// in: captured graphs (one per GMRES cycle body) and the fused launches'
// tables.  Bounded: the oldest entry goes first.
struct LaunchCache {
    struct Entry {
        std::vector<uintptr_t> sig;
        hipGraphExec_t exec = nullptr;
        std::unique_ptr<FusePlan> plan;
    };
    std::vector<std::unique_ptr<Entry>> entries;
    hipStream_t capture = nullptr;  // non-default stream used only for capture
    size_t limit = 8;
    LaunchCache() = default;
    LaunchCache(const LaunchCache &) = delete;
    LaunchCache &operator=(const LaunchCache &) = delete;
    ~LaunchCache() {
        clear();
        if (capture) (void)hipStreamDestroy(capture);
    }
    void clear() {
        for (auto &e : entries)
            if (e->exec) (void)hipGraphExecDestroy(e->exec);
        entries.clear();
    }
    Entry *find(const std::vector<uintptr_t> &sig) {
        for (auto &e : entries)
            if (e->sig == sig) return e.get();
        return nullptr;
    }
    Entry *add(std::vector<uintptr_t> sig) {
        if (entries.size() >= limit) {
            if (entries.front()->exec) (void)hipGraphExecDestroy(entries.front()->exec);
            entries.erase(entries.begin());
        }
        entries.push_back(std::make_unique<Entry>());
        entries.back()->sig = std::move(sig);
        return entries.back().get();
    }
};

}  // namespace

struct hz_operator {
    int nx = 0, ny = 0, nz = 0;
    int64_t n = 0;
    double dx = 0, omega = 0;
    int soft = 0;
    int open_mask = 0;  // bit f: face f is open (radiating)
    DevBuf rho64, kap64, rho32, kap32;
    DevBuf a64[3], q64[3], a32[3], q32[3];
    DevBuf face64[6], face32[6];
    DevBuf rsf64, mf64, rsa64, ma64, rsf32, mf32, rsa32, ma32;
    OpView<double> v64{};
    OpView<float> v32{};
    TileMap tm{};
    int slab_shape = 0;  // slab tile shape of the complex64 apply
    // Source columns, grouped by node (scatter) and by column (gather).
    int ncols = 0;
    int64_t nrows_src = 0;
    DevBuf src_row_node, src_row_ptr, src_row_col, src_row_val;
    DevBuf src_col_node, src_col_ptr, src_col_val;
    // Cached GMRES workspace (grow-only).
    DevBuf V, X, Bh, dpart, npart, cpart, red, Hm, gc, gs, g, scales, coef, resid, tol, steps, live, member;
    DevBuf Hu, coef2, alpha, gram, bfix;  // two-step Arnoldi blocks
    DevBuf Z;                       // multigrid form: preconditioned basis slots, restart x k x n
    DevBuf rhs32;
    std::unique_ptr<Multigrid> mg;
    LaunchCache graphs;  // captured GMRES cycle bodies of this operator's solves
    hz_operator() = default;
    hz_operator(const hz_operator &) = delete;
    hz_operator &operator=(const hz_operator &) = delete;
};

namespace {

template <typename T>
const OpView<T> &view(const hz_operator *op);
template <>
const OpView<double> &view<double>(const hz_operator *op) { return op->v64; }
template <>
const OpView<float> &view<float>(const hz_operator *op) { return op->v32; }

int64_t vec_blocks(int64_t n) { return blocks_for(n, 256 * 16, 1024); }
// Workgroups per column of the Krylov vector passes over k columns: those of
// vec_blocks, except that a single column gets twice as many (at 2^21
// complex128 elements a 16-vector dot pass streams 6.1 TB/s on 1024
// workgroups and 4.0 TB/s on 512).
int64_t vec_blocks_cols(int64_t n, int k) { return k == 1 ? blocks_for(n, 256 * 8, 1024) : vec_blocks(n); }
// Elements per workgroup of the vector kernels, even so that paired loads
// never straddle two workgroups.
int64_t vec_chunk(int64_t n, int64_t nblk) { return ((n + nblk - 1) / nblk + 1) & ~int64_t(1); }
// Paired 16-byte loads for complex64 vectors of even length.
template <typename T>
bool paired(int64_t n) {
    return std::is_same<T, float>::value && n % 2 == 0;
}

// Vector width of the Gram-Schmidt passes over k columns of n reading
// `vecs` basis vectors (see STREAM_MIN_BYTES): 3 streamed pairs, 2 cached
// pairs, 4 single values with the basis streamed (complex128 only), 1
// single values.
template <typename T>
int vec_width(int64_t n, int k, int vecs = 0) {
    const int64_t bytes = (int64_t)k * n * (int64_t)sizeof(cx<T>);
    if (!paired<T>(n)) return std::is_same<T, double>::value && vecs * bytes >= STREAM_PASS_BYTES ? 4 : 1;
    return bytes >= STREAM_MIN_BYTES ? 3 : 2;
}

// ------------------------------------------------------------- fused launches
// The solver's host code sends its kernel launches to a Fuser.  Without a
// plan it launches each kernel at once (one scene).  In a lockstep batch it
// collects the launches the scenes make at one point of the sequence and
// flush() issues them: one k_multi launch per distinct kernel (hz_multi.hpp),
// or the kernel's own launch where a single scene makes it.  A scene that
// launches again before a flush flushes first, so a fused launch holds at
// most one launch per scene and every scene's own launch order is kept.
// Launch counters (hz_launch_counts): graph launches, fused (k_multi)
// kernel launches and single-scene kernel launches made on a stream; a
// sequence being captured counts once per graph launch instead.
std::atomic<int64_t> g_launches[3];

template <class K>
void launch_multi(const void *table, const MultiGrid &mg, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_multi<K>), dim3(blocks), dim3(K::THREADS), 0, s,
                       static_cast<const typename K::Args *>(table), mg);
}
template <class K>
void launch_single(dim3 grid, hipStream_t s, const void *args) {
    K::single(grid, s, *static_cast<const typename K::Args *>(args));
}

class Fuser {
  public:
    // counted: launches count in the launch counters (not under capture).
    explicit Fuser(hipStream_t s, FusePlan *plan = nullptr, bool counted = true)
        : s_(s), plan_(plan), counted_(counted) {}
    Fuser(const Fuser &) = delete;
    Fuser &operator=(const Fuser &) = delete;
    hipStream_t stream() const { return s_; }

    template <class K>
    void add(int scene, dim3 grid, const typename K::Args &a) {
        if (!plan_) {
            K::single(grid, s_, a);
            HZ_LAUNCHED();
            if (counted_) ++g_launches[2];
            return;
        }
        static_assert(std::is_trivially_copyable<typename K::Args>::value, "launch arguments are copied as bytes");
        if (std::find(pending_.begin(), pending_.end(), scene) != pending_.end()) flush();
        pending_.push_back(scene);
        const void *key = reinterpret_cast<const void *>(&launch_multi<K>);
        Group *g = nullptr;
        for (Group &x : groups_)
            if (x.key == key) g = &x;
        if (!g) {
            groups_.push_back(Group{key, sizeof(typename K::Args), &launch_multi<K>, &launch_single<K>, {}, {}});
            g = &groups_.back();
        }
        const char *p = reinterpret_cast<const char *>(&a);
        g->args.insert(g->args.end(), p, p + sizeof(a));
        g->grids.push_back(grid);
    }

    // Device-to-device copy, in launch order.
    void copy(void *dst, const void *src, size_t bytes) {
        flush();
        if (plan_ && plan_->recording) return;
        HZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s_));
    }

    void flush() {
        pending_.clear();
        if (groups_.empty()) return;
        std::vector<Group> gs;
        gs.swap(groups_);
        const bool rec = plan_->recording;
        for (Group &g : gs) {
            const int count = (int)g.grids.size();
            if (count == 1) {
                if (!rec) {
                    g.one(g.grids[0], s_, g.args.data());
                    HZ_LAUNCHED();
                    if (counted_) ++g_launches[2];
                }
                continue;
            }
            for (int c0 = 0; c0 < count; c0 += MULTI_MAX) {
                const int cnt = std::min(MULTI_MAX, count - c0);
                MultiGrid mg{};
                mg.count = cnt;
                int64_t total = 0;
                for (int i = 0; i < cnt; ++i) {
                    const dim3 gd = g.grids[c0 + i];
                    mg.first[i] = (int)total;
                    mg.gx[i] = (int)gd.x;
                    mg.gy[i] = (int)gd.y;
                    mg.gz[i] = (int)gd.z;
                    total += ((int64_t)gd.x * gd.y * gd.z + 7) / 8 * 8;
                }
                if (total >= (int64_t(1) << 31)) throw BadArgument("fused launch too large");
                for (int i = cnt; i <= MULTI_MAX; ++i) mg.first[i] = (int)total;
                const size_t bytes = (size_t)cnt * g.size;
                if (rec) {
                    const size_t off = (plan_->host.size() + 255) / 256 * 256;
                    plan_->host.resize(off + bytes);
                    std::memcpy(plan_->host.data() + off, g.args.data() + (size_t)c0 * g.size, bytes);
                    plan_->offs.push_back(off);
                } else {
                    if (plan_->next >= plan_->offs.size())
                        throw std::logic_error("fused launch sequence differs from its recording");
                    g.multi(plan_->dev.as<char>() + plan_->offs[plan_->next++], mg, (unsigned)total, s_);
                    HZ_LAUNCHED();
                    if (counted_) ++g_launches[1];
                }
            }
        }
    }

  private:
    struct Group {
        const void *key;
        size_t size;
        void (*multi)(const void *, const MultiGrid &, unsigned, hipStream_t);
        void (*one)(dim3, hipStream_t, const void *);
        std::vector<char> args;
        std::vector<dim3> grids;
    };
    hipStream_t s_;
    FusePlan *plan_;
    bool counted_;
    std::vector<Group> groups_;
    std::vector<int> pending_;
};

// ---------------------------------------------------------------- launchers
// Each takes the Fuser and the scene index of the launch.
template <typename T>
void launch_dots(Fuser &f, int sc, dim3 grid, const cx<T> *V, int k, int64_t n, int jp1, const cx<T> *w,
                 const int *live, double2 *part, int maxv, int64_t chunk) {
    const int vw = vec_width<T>(n, k, jp1);
    if constexpr (std::is_same<T, float>::value) {
        if (vw == 3) {
            f.add<KDots<T, 3>>(sc, grid, {V, k, n, jp1, w, live, part, maxv, chunk});
            return;
        }
        if (vw == 2) {
            f.add<KDots<T, 2>>(sc, grid, {V, k, n, jp1, w, live, part, maxv, chunk});
            return;
        }
    } else {
        if (vw == 4) {
            f.add<KDots<T, 4>>(sc, grid, {V, k, n, jp1, w, live, part, maxv, chunk});
            return;
        }
    }
    f.add<KDots<T, 1>>(sc, grid, {V, k, n, jp1, w, live, part, maxv, chunk});
}

template <typename T>
void launch_update(Fuser &f, int sc, dim3 grid, const cx<T> *V, int k, int64_t n, int cnt_all, const int *cnt_rhs,
                   cx<T> *target, int64_t ldt, const cx<double> *coef, int maxv, const int *live, double *norm_partial,
                   int64_t chunk) {
    // The cycle-end update (counts per column) reads at most maxv - 1
    // vectors; paired accesses need an even target leading dimension too.
    int vw = vec_width<T>(n, k, cnt_rhs ? maxv - 1 : cnt_all);
    if ((vw == 2 || vw == 3) && ldt % 2 != 0) vw = 1;
    if constexpr (std::is_same<T, float>::value) {
        if (vw == 3) {
            f.add<KUpdate<T, 3>>(sc, grid, {V, k, n, cnt_all, cnt_rhs, target, ldt, coef, maxv, live, norm_partial,
                                            chunk});
            return;
        }
        if (vw == 2) {
            f.add<KUpdate<T, 2>>(sc, grid, {V, k, n, cnt_all, cnt_rhs, target, ldt, coef, maxv, live, norm_partial,
                                            chunk});
            return;
        }
    } else {
        if (vw == 4) {
            f.add<KUpdate<T, 4>>(sc, grid, {V, k, n, cnt_all, cnt_rhs, target, ldt, coef, maxv, live, norm_partial,
                                            chunk});
            return;
        }
    }
    f.add<KUpdate<T, 1>>(sc, grid, {V, k, n, cnt_all, cnt_rhs, target, ldt, coef, maxv, live, norm_partial, chunk});
}

// Workgroups of the per-node multigrid transfers: one per 256 nodes of an
// xy-plane, z-plane and slot, padded to a multiple of 8 for the XCD remap.
int plane_blocks(int nx, int ny, int nz, int k) {
    const int64_t t = (int64_t)((nx * ny + 255) / 256) * nz * k;
    return (int)((t + 7) / 8 * 8);
}

// Slots per thread of the per-coarse-node transfers: as many as the grid
// allows while it keeps >= 2048 workgroups.
int transfer_group(const MgLevel &C, int k) {
    const int per_slot = plane_blocks(C.nx, C.ny, C.nz, 1);
    const int groups = std::max(1, std::min(k, (2048 + per_slot - 1) / per_slot));
    return (k + groups - 1) / groups;
}

void launch_restrict(Fuser &f, int sc, const MgLevel &L, const MgLevel &C, const cx<float> *r, cx<float> *bc, int k,
                     const int *live) {
    // z-marching restriction: one workgroup per 256 coarse nodes of a plane,
    // chunk of RZ_CHUNK coarse planes and slot.
    const int64_t t = (int64_t)((C.nx * C.ny + 255) / 256) * ((C.nz + RZ_CHUNK - 1) / RZ_CHUNK) * k;
    const int nb = (int)((t + 7) / 8 * 8);
    if (L.nx % 2 == 0)
        f.add<KRestrictZ<float, 1>>(sc, dim3(nb), {r, L.nx, L.ny, L.nz, bc, C.nx, C.ny, C.nz, k, live, L.v.soft, nb});
    else
        f.add<KRestrictZ<float, 0>>(sc, dim3(nb), {r, L.nx, L.ny, L.nz, bc, C.nx, C.ny, C.nz, k, live, L.v.soft, nb});
}

void launch_prolong(Fuser &f, int sc, const MgLevel &L, const MgLevel &C, const cx<float> *xc, cx<float> *xf, int k,
                    const int *live, const cx<float> *b, const cx<float> *dinv, float w, const float *bdiv,
                    int conj_d) {
    // One thread per coarse cell and group of rs slots: the slots of a group
    // share the cell's weights and first-sweep factors.  At most 8 slots per
    // group (while the grid keeps >= 2048 workgroups): the 128^3 -> 65^3
    // prolongation of 64 columns inside a solve takes 531 us with groups of
    // 8, 574 us with 4 and 583 us with 32 (the grid's limit there).
    const int rs = std::min(8, transfer_group(C, k));
    const int nb = plane_blocks(C.nx, C.ny, C.nz, (k + rs - 1) / rs);
    if (L.nx % 2 == 0)
        f.add<KProlong<float, 1>>(sc, dim3(nb), {xc, C.nx, C.ny, C.nz, xf, L.nx, L.ny, L.nz, k, rs, live, L.v.soft, b,
                                                 dinv, w, bdiv, conj_d, nb});
    else
        f.add<KProlong<float, 0>>(sc, dim3(nb), {xc, C.nx, C.ny, C.nz, xf, L.nx, L.ny, L.nz, k, rs, live, L.v.soft, b,
                                                 dinv, w, bdiv, conj_d, nb});
}

// One slab-stencil launch with tile shape `shape` (index into SLAB_SHAPE).
template <int MODE, int EPI>
void launch_slab(Fuser &f, int sc, int shape, const OpView<float> &v, const cx<float> *x, int64_t ldx, cx<float> *y,
                 int64_t ldy, int k, const int *live, const double *scale, const cx<float> *b, const cx<float> *dinv,
                 float w, const float *bdiv) {
    // Paired tiles move two nodes per 16-byte access: even widths and
    // 16-byte aligned columns only.
    auto a16 = [](const void *ptr) { return ((uintptr_t)ptr & 15) == 0; };
    if (SLAB_SHAPE[shape].flat == 4 &&
        !(v.nx % 2 == 0 && ldx % 2 == 0 && ldy % 2 == 0 && a16(x) && a16(y) && a16(b)))
        shape = PAIR_FALLBACK;
    const TileMap tm = slab_map(v.nx, v.ny, v.nz, shape, k);
    const dim3 grid(tm.nblocks * tm.csplit);
#define HZ_SLAB_CASE(ID, ZB, TY, RP, NT, FLAT)                                                                     \
    case ID:                                                                                                       \
        static_assert(SLAB_SHAPE[ID].zb == ZB && SLAB_SHAPE[ID].ty == TY && SLAB_SHAPE[ID].rp == RP &&             \
                          SLAB_SHAPE[ID].nt == NT && SLAB_SHAPE[ID].flat == FLAT,                                  \
                      "slab shape table");                                                                         \
        f.add<KSlab<MODE, EPI, ZB, TY, RP, NT, FLAT>>(sc, grid, {v, tm, x, ldx, y, ldy, k, live, scale, b, dinv, w, \
                                                                 bdiv});                                           \
        break;
    switch (shape) {
        HZ_SLAB_CASE(0, 4, 4, 1, 0, 0)
        HZ_SLAB_CASE(1, 2, 4, 1, 0, 0)
        HZ_SLAB_CASE(2, 4, 8, 1, 0, 0)
        HZ_SLAB_CASE(3, 2, 8, 1, 0, 0)
        HZ_SLAB_CASE(4, 4, 2, 1, 0, 0)
        HZ_SLAB_CASE(5, 2, 4, 2, 0, 0)
        HZ_SLAB_CASE(6, 4, 4, 1, 1, 0)
        HZ_SLAB_CASE(7, 2, 4, 2, 1, 0)
        HZ_SLAB_CASE(8, 4, 4, 1, 1, 1)
        HZ_SLAB_CASE(9, 2, 4, 1, 1, 1)
        HZ_SLAB_CASE(10, 2, 4, 2, 1, 1)
        HZ_SLAB_CASE(11, 4, 4, 1, 1, 2)
        HZ_SLAB_CASE(12, 2, 4, 2, 1, 2)
        HZ_SLAB_CASE(13, 4, 8, 1, 1, 2)
#define HZ_SLAB2_CASE(ID, ZB, TY, NT)                                                                              \
    case ID:                                                                                                       \
        static_assert(SLAB_SHAPE[ID].zb == ZB && SLAB_SHAPE[ID].ty == TY && SLAB_SHAPE[ID].rp == 1 &&              \
                          SLAB_SHAPE[ID].nt == NT && SLAB_SHAPE[ID].flat == 4,                                     \
                      "slab shape table");                                                                         \
        f.add<KSlab2<MODE, EPI, ZB, TY, NT>>(sc, grid, {v, tm, x, ldx, y, ldy, k, live, scale, b, dinv, w, bdiv});  \
        break;
        HZ_SLAB2_CASE(14, 2, 4, 1)
        HZ_SLAB2_CASE(15, 2, 8, 1)
        HZ_SLAB2_CASE(16, 4, 4, 1)
#undef HZ_SLAB2_CASE
        default:
            throw BadArgument("unknown slab tile shape");
    }
#undef HZ_SLAB_CASE
}

template <typename T, int MODE>
void launch_apply(Fuser &f, int sc, const hz_operator *op, const cx<T> *x, int64_t ldx, cx<T> *y, int64_t ldy,
                  int nrhs, const int *live, const double *scale) {
    for (int r0 = 0; r0 < nrhs; r0 += APPLY_MAX_RHS) {
        const int kk = std::min(APPLY_MAX_RHS, nrhs - r0);
        const cx<T> *xs = x + (int64_t)r0 * ldx;
        cx<T> *ys = y + (int64_t)r0 * ldy;
        const int *lv = live ? live + r0 : nullptr;
        const double *sc_r = scale ? scale + r0 : nullptr;
        if constexpr (std::is_same<T, float>::value) {
            // complex64: the 2.5-D slab kernel (register z-window, DPP x-neighbours)
            launch_slab<MODE, SLAB_APPLY>(f, sc, op->slab_shape, op->v32, xs, ldx, ys, ldy, kk, lv, sc_r, nullptr,
                                          nullptr, 0.0f, nullptr);
        } else {
            // complex128: the same slab layout with FP64 row coefficients
            const TileMap tm = slab64_map(op->nx, op->ny, op->nz, kk);
            f.add<KSlab64<MODE, SLAB64_ZB, SLAB64_TY>>(sc, dim3(tm.nblocks * tm.csplit),
                                                       {view<T>(op), tm, xs, ldx, ys, ldy, kk, lv, sc_r});
        }
    }
}

template <typename T>
void apply_mode(Fuser &f, int sc, const hz_operator *op, int mode, const cx<T> *x, int64_t ldx, cx<T> *y,
                int64_t ldy, int nrhs, const int *live, const double *scale) {
    switch (mode) {
        case MODE_A: launch_apply<T, MODE_A>(f, sc, op, x, ldx, y, ldy, nrhs, live, scale); break;
        case MODE_RAM: launch_apply<T, MODE_RAM>(f, sc, op, x, ldx, y, ldy, nrhs, live, scale); break;
        case MODE_AH: launch_apply<T, MODE_AH>(f, sc, op, x, ldx, y, ldy, nrhs, live, scale); break;
        case MODE_RAMH: launch_apply<T, MODE_RAMH>(f, sc, op, x, ldx, y, ldy, nrhs, live, scale); break;
        case MODE_RA: launch_apply<T, MODE_RA>(f, sc, op, x, ldx, y, ldy, nrhs, live, scale); break;
        case MODE_RAH: launch_apply<T, MODE_RAH>(f, sc, op, x, ldx, y, ldy, nrhs, live, scale); break;
        default: throw BadArgument("unknown apply mode");
    }
}

template <typename T>
void residual_mode(Fuser &f, int sc, const hz_operator *op, int mode, const cx<T> *x, const cx<T> *b, cx<T> *out,
                   int k, const int *live, double *partial) {
    const dim3 grid(op->tm.nblocks, k);
    const int64_t n = op->n;
    // complex64: the slab stencil's residual epilogue (b - op(x), the row in
    // the order of k_residual), whose norm partials col_norm_partials adds;
    // complex128: k_residual with its partials over op->tm.
    auto go = [&](auto mode_c) {
        constexpr int M = decltype(mode_c)::value;
        if constexpr (std::is_same<T, float>::value)
            launch_slab<M, SLAB_RESID>(f, sc, op->slab_shape, op->v32, x, n, out, n, k, live, nullptr, b, nullptr,
                                       0.0f, nullptr);
        else
            f.add<KResidual<T, M>>(sc, grid, {view<T>(op), op->tm, x, n, b, n, out, n, live, partial});
    };
    switch (mode) {
        case MODE_RAM: go(std::integral_constant<int, MODE_RAM>()); break;
        case MODE_RAMH: go(std::integral_constant<int, MODE_RAMH>()); break;
        case MODE_RA: go(std::integral_constant<int, MODE_RA>()); break;
        case MODE_RAH: go(std::integral_constant<int, MODE_RAH>()); break;
        default: throw BadArgument("unknown residual mode");
    }
}

// FP64 norm partials of k complex64 columns of n (leading dimension n) over
// nblk chunks of `chunk` elements: partial[r * nblk + blk].
void col_norm_partials(Fuser &f, int sc, const cx<float> *v, int64_t n, int k, const int *live, double *partial,
                       int64_t nblk, int64_t chunk) {
    const dim3 grid((unsigned)nblk, k);
    const int vw = vec_width<float>(n, k);
    if (vw == 3)
        f.add<KColNorms<3>>(sc, grid, {v, n, n, live, partial, chunk});
    else if (vw == 2)
        f.add<KColNorms<2>>(sc, grid, {v, n, n, live, partial, chunk});
    else
        f.add<KColNorms<1>>(sc, grid, {v, n, n, live, partial, chunk});
}

// ------------------------------------------------------------------ multigrid
void mg_reserve(Multigrid &mg, int k) {
    if (k <= mg.kcap) return;
    for (auto &L : mg.levels) {
        const size_t bytes = (size_t)k * L->n * sizeof(cx<float>);
        L->bb.ensure(bytes);
        L->xx.ensure(bytes);
        L->tt.ensure(bytes);
    }
    mg.kcap = k;
}

template <int MODE>
void level_smooth(Fuser &f, int sc, MgLevel &L, const cx<float> *x, const cx<float> *b, cx<float> *out, float w,
                  int k, const int *live, const float *bdiv) {
    launch_slab<MODE, SLAB_SMOOTH>(f, sc, L.shape[0], L.v, x, L.n, out, L.n, k, live, nullptr, b,
                                   L.dinv.as<cx<float>>(), w, bdiv);
}

// One scene's V-cycle: x = V-cycle(b / bdiv) on level l of mg; x and b hold k
// slots of length n_l (x must not alias b).  bdiv (finest level only) is the
// Krylov system's row scale.
struct MgTask {
    int scene;
    Multigrid *mg;
    const cx<float> *b;
    cx<float> *x;
    int k;
    const int *live;
    const float *bdiv;
};

// The V-cycles of several scenes on level l, in lockstep: every launch
// position is one flush of the Fuser, so the scenes' launches there fuse.
// A scene whose hierarchy ends above level l+1 takes its coarsest-level
// path while the others descend.
struct VcState {
    MgTask t;
    MgLevel *L, *C;
    bool coarsest, direct, pre1;
    int post, first;
    cx<float> *cur, *alt;
};

template <int MODE>
void vcycle(Fuser &f, std::vector<MgTask> &tasks, int l) {
    using State = VcState;
    std::vector<State> st;
    std::vector<MgTask> down;
    for (const MgTask &t : tasks) {
        State s{};
        s.t = t;
        Multigrid &mg = *t.mg;
        s.L = mg.levels[l].get();
        s.coarsest = (l + 1 == (int)mg.levels.size());
        s.C = s.coarsest ? nullptr : mg.levels[l + 1].get();
        s.direct = s.coarsest && s.L->n <= COARSE_MAX_NODES && !t.bdiv;
        s.pre1 = !s.coarsest && mg.pre == 1;
        s.post = l == 0 ? mg.post : mg.post_coarse;
        st.push_back(s);
    }
    // Pre-smoothing: the coarsest level's direct solve; one pre-sweep folded
    // into the residual (its iterate f b, f = w dinv / bdiv, is never
    // stored: the residual b/bdiv - A (f b) comes from b in one stencil pass
    // and the prolongation rebuilds f b + P e from b); or explicit Jacobi
    // sweeps, ping-ponging between x and the level scratch so that the last
    // one lands in x.
    for (State &s : st) {
        MgLevel &L = *s.L;
        const MgTask &t = s.t;
        const cx<float> *dinv = L.dinv.as<cx<float>>();
        if (s.direct) {
            f.add<KCoarseJacobi<float, MODE>>(t.scene, dim3(t.k), {L.v, t.b, t.x, dinv, t.mg->weight,
                                                                   t.mg->coarse_sweeps, t.live});
        } else if (s.pre1) {
            launch_slab<MODE, SLAB_RESID0>(f, t.scene, L.shape[2], L.v, t.b, L.n, L.tt.as<cx<float>>(), L.n, t.k,
                                           t.live, nullptr, nullptr, dinv, t.mg->weight, t.bdiv);
        } else {
            const int swaps = s.coarsest ? t.mg->coarse_sweeps - 1 : (t.mg->pre - 1) + s.post;
            s.cur = (swaps % 2 == 0) ? t.x : L.tt.as<cx<float>>();
            s.alt = (s.cur == t.x) ? L.tt.as<cx<float>>() : t.x;
            s.first = s.coarsest ? t.mg->coarse_sweeps : t.mg->pre;
            f.add<KLevelJacobi0<float, MODE>>(t.scene, dim3(blocks_for(L.n, 256 * 8, 2048), t.k),
                                              {t.b, s.cur, dinv, t.mg->weight, L.n, t.live, t.bdiv});
        }
    }
    f.flush();
    for (int it = 1;; ++it) {
        bool any = false;
        for (State &s : st) {
            if (s.direct || s.pre1 || it >= s.first) continue;
            level_smooth<MODE>(f, s.t.scene, *s.L, s.cur, s.t.b, s.alt, s.t.mg->weight, s.t.k, s.t.live, s.t.bdiv);
            std::swap(s.cur, s.alt);
            any = true;
        }
        f.flush();
        if (!any) break;
    }
    // Coarse-grid correction.
    for (State &s : st) {
        if (s.coarsest || s.pre1) continue;
        launch_slab<MODE, SLAB_RESID>(f, s.t.scene, s.L->shape[1], s.L->v, s.cur, s.L->n, s.alt, s.L->n, s.t.k,
                                      s.t.live, nullptr, s.t.b, nullptr, 0.0f, s.t.bdiv);
    }
    f.flush();
    for (State &s : st) {
        if (s.coarsest) continue;
        const cx<float> *r = s.pre1 ? s.L->tt.as<cx<float>>() : s.alt;
        launch_restrict(f, s.t.scene, *s.L, *s.C, r, s.C->bb.as<cx<float>>(), s.t.k, s.t.live);
        down.push_back({s.t.scene, s.t.mg, s.C->bb.as<cx<float>>(), s.C->xx.as<cx<float>>(), s.t.k, s.t.live,
                        nullptr});
    }
    f.flush();
    if (!down.empty()) vcycle<MODE>(f, down, l + 1);
    for (State &s : st) {
        if (s.coarsest) continue;
        const cx<float> *dinv = s.L->dinv.as<cx<float>>();
        if (s.pre1) {
            // Post-sweeps start where the last one lands in x.
            s.cur = (s.post % 2 == 0) ? s.t.x : s.L->tt.as<cx<float>>();
            s.alt = (s.cur == s.t.x) ? s.L->tt.as<cx<float>>() : s.t.x;
            launch_prolong(f, s.t.scene, *s.L, *s.C, s.C->xx.as<cx<float>>(), s.cur, s.t.k, s.t.live, s.t.b, dinv,
                           s.t.mg->weight, s.t.bdiv, (int)(MODE == MODE_AH));
        } else {
            launch_prolong(f, s.t.scene, *s.L, *s.C, s.C->xx.as<cx<float>>(), s.cur, s.t.k, s.t.live, nullptr, dinv,
                           s.t.mg->weight, s.t.bdiv, 0);
        }
    }
    f.flush();
    for (int it = 0;; ++it) {
        bool any = false;
        for (State &s : st) {
            if (s.coarsest || it >= s.post) continue;
            level_smooth<MODE>(f, s.t.scene, *s.L, s.cur, s.t.b, s.alt, s.t.mg->weight, s.t.k, s.t.live, s.t.bdiv);
            std::swap(s.cur, s.alt);
            any = true;
        }
        f.flush();
        if (!any) break;
    }
    // W-cycle below the finest level: further coarse-grid corrections of x
    // (one-pre-sweep levels).
    for (int visit = 1;; ++visit) {
        std::vector<State *> w;
        for (State &s : st)
            if (l >= 1 && s.pre1 && visit < s.t.mg->cycle) w.push_back(&s);
        if (w.empty()) break;
        for (State *s : w)
            launch_slab<MODE, SLAB_RESID>(f, s->t.scene, s->L->shape[1], s->L->v, s->t.x, s->L->n,
                                          s->L->tt.as<cx<float>>(), s->L->n, s->t.k, s->t.live, nullptr, s->t.b,
                                          nullptr, 0.0f, s->t.bdiv);
        f.flush();
        std::vector<MgTask> again;
        for (State *s : w) {
            launch_restrict(f, s->t.scene, *s->L, *s->C, s->L->tt.as<cx<float>>(), s->C->bb.as<cx<float>>(), s->t.k,
                            s->t.live);
            again.push_back({s->t.scene, s->t.mg, s->C->bb.as<cx<float>>(), s->C->xx.as<cx<float>>(), s->t.k,
                             s->t.live, nullptr});
        }
        f.flush();
        vcycle<MODE>(f, again, l + 1);
        for (State *s : w) {
            launch_prolong(f, s->t.scene, *s->L, *s->C, s->C->xx.as<cx<float>>(), s->t.x, s->t.k, s->t.live, nullptr,
                           s->L->dinv.as<cx<float>>(), s->t.mg->weight, s->t.bdiv, 0);
            s->cur = s->t.x;
            s->alt = s->L->tt.as<cx<float>>();
        }
        f.flush();
        for (int it = 0;; ++it) {
            bool any = false;
            for (State *s : w) {
                if (it >= s->post) continue;
                level_smooth<MODE>(f, s->t.scene, *s->L, s->cur, s->t.b, s->alt, s->t.mg->weight, s->t.k, s->t.live,
                                   s->t.bdiv);
                std::swap(s->cur, s->alt);
                any = true;
            }
            f.flush();
            if (!any) break;
        }
        for (State *s : w)
            if (s->cur != s->t.x) f.copy(s->t.x, s->cur, (size_t)s->t.k * s->L->n * sizeof(cx<float>));
    }
    // Explicit-sweep levels end in their scratch when the sweep count is odd.
    for (State &s : st)
        if (!s.direct && !s.pre1 && s.cur != s.t.x)
            f.copy(s.t.x, s.cur, (size_t)s.t.k * s.L->n * sizeof(cx<float>));
}

// z = P R^-1 v: one V-cycle of each scene on its unscaled right-hand side.
void precondition(Fuser &f, std::vector<MgTask> &tasks, bool adjoint) {
    if (adjoint)
        vcycle<MODE_AH>(f, tasks, 0);
    else
        vcycle<MODE_A>(f, tasks, 0);
}

// Everything a captured cycle bakes into its launches: precision, mode, chunk
// width, restart, preconditioner, tile shapes and every buffer address.
std::vector<uintptr_t> graph_signature(const hz_operator *op, size_t elem, int mode, int k, int restart,
                                       bool use_mg) {
    std::vector<uintptr_t> sig = {elem, (uintptr_t)mode, (uintptr_t)k, (uintptr_t)restart, (uintptr_t)use_mg,
                                  (uintptr_t)op->slab_shape};
    for (const DevBuf *b : {&op->V, &op->X, &op->Z, &op->dpart, &op->npart, &op->Hm, &op->gc, &op->gs, &op->g,
                            &op->scales, &op->coef, &op->resid, &op->tol, &op->steps, &op->live, &op->member,
                            &op->Hu, &op->coef2, &op->alpha, &op->bfix, &op->gram, &op->Bh})
        sig.push_back((uintptr_t)b->p);
    if (use_mg)
        for (const auto &L : op->mg->levels) {
            for (const DevBuf *b : {&L->bb, &L->xx, &L->tt}) sig.push_back((uintptr_t)b->p);
            for (int e = 0; e < 3; ++e) sig.push_back((uintptr_t)L->shape[e]);
        }
    return sig;
}

// Runs the launch sequence `body(Fuser &)` on stream s.  Without fusion or
// capture it launches as it goes.  Otherwise the sequence is kept in `lc`
// under `sig`: fused sequences are recorded once (their tables uploaded),
// captured ones become a graph; later runs replay them.  A capture records
// launches only (no memset nodes).
template <typename F>
void run_sequence(LaunchCache &lc, std::vector<uintptr_t> sig, bool fused, bool graph, hipStream_t s, F &&body) {
    if (!fused && !graph) {
        Fuser f(s);
        body(f);
        return;
    }
    LaunchCache::Entry *e = lc.find(sig);
    if (!e) {
        e = lc.add(std::move(sig));
        if (fused) {
            e->plan = std::make_unique<FusePlan>();
            Fuser rec(s, e->plan.get());
            body(rec);
            rec.flush();
            e->plan->upload(s);
        }
        if (graph) {
            if (!lc.capture) HZ_HIP(hipStreamCreateWithFlags(&lc.capture, hipStreamNonBlocking));
            HZ_HIP(hipStreamBeginCapture(lc.capture, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                if (e->plan) e->plan->next = 0;
                Fuser f(lc.capture, e->plan.get(), false);
                body(f);
                f.flush();
            } catch (...) {
                (void)hipStreamEndCapture(lc.capture, &g);
                if (g) (void)hipGraphDestroy(g);
                lc.entries.pop_back();
                throw;
            }
            HZ_HIP(hipStreamEndCapture(lc.capture, &g));
            hipGraphExec_t exec = nullptr;
            const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) lc.entries.pop_back();
            HZ_HIP(err);
            e->exec = exec;
        }
    }
    if (e->exec) {
        HZ_HIP(hipGraphLaunch(e->exec, s));
        ++g_launches[0];
        return;
    }
    e->plan->next = 0;
    Fuser f(s, e->plan.get());
    body(f);
    f.flush();
}

// Per-RHS bookkeeping of one chunk.
struct ChunkState {
    int k = 0;
    std::vector<int> active;      // still being solved
    std::vector<int> iterations;  // inner steps
    std::vector<int> cycles;
    std::vector<double> last_resid;
};

// Preconditioner of the inner solve: Jacobi (operator R A M, the reference's
// choice) or shifted-Laplacian multigrid P (operator R A P).
enum Precond { PRE_JACOBI = 0, PRE_MULTIGRID = 1 };

// One scene's chunk in a GMRES run: its operator (whose workspace holds the
// chunk), its k columns' progress and their absolute tolerances.
struct Member {
    hz_operator *op = nullptr;
    int k = 0;
    ChunkState *cs = nullptr;
    std::vector<double> atol;
};

// Device view of an operator's GMRES state.
GmresState gmres_state(hz_operator *op) {
    GmresState st;
    st.H = op->Hm.as<cx<double>>();
    st.gc = op->gc.as<double>();
    st.gs = op->gs.as<cx<double>>();
    st.g = op->g.as<cx<double>>();
    st.scales = op->scales.as<double>();
    st.coef = op->coef.as<cx<double>>();
    st.resid = op->resid.as<double>();
    st.tol = op->tol.as<double>();
    st.steps = op->steps.as<int>();
    st.live = op->live.as<int>();
    st.Hu = op->Hu.as<cx<double>>();
    st.coef2 = op->coef2.as<cx<double>>();
    st.alpha = op->alpha.as<cx<double>>();
    st.fix = op->bfix.as<int>();
    st.guard = block_guard();
    return st;
}

// Restarted GMRES on the k slots of every member, (R A M) y = Bh in
// precision T with y in the member's X.  With several members (scenes) the
// cycles run in lockstep: every launch position fuses across the scenes that
// are still iterating, and each cycle body is one graph launch for all of
// them.
//   ref_semantics: full `restart` steps per cycle, convergence tested only at
//   the cycle top (the reference/CuPy loop); otherwise a slot leaves its cycle
//   as soon as the Givens estimate meets its tolerance.
// On return X holds y (M y in the multigrid form) for every slot that was
// active on entry.
template <typename T>
void gmres_cycles(std::vector<Member> &ms, int mode_ram, int restart, int maxiter, bool ref_semantics,
                  hipStream_t s, int precond = PRE_JACOBI, bool use_graph = false, int block = 1,
                  LaunchCache *batch_cache = nullptr) {
    const int S = (int)ms.size();
    const bool fused = S > 1;
    const bool use_mg = precond == PRE_MULTIGRID;
    const bool adjoint = (mode_ram == MODE_RAMH);
    const int mode_op = use_mg ? (adjoint ? MODE_RAH : MODE_RA) : mode_ram;
    const int maxv = restart + 1;
    const bool two = block == 2 && std::is_same<T, float>::value;
    struct Run {
        hz_operator *op;
        int k;
        int64_t n, nblk, chunk;
        cx<T> *V, *X, *Bh;
        cx<float> *Z;
        GmresState st;
        int *member;
        double2 *dpart;  // dot-product partials
        double *npart;   // norm partials
        double *gram;    // two-step block Gram partials
        std::vector<double> early, resid, prev_top;
        std::vector<int> live_h, early_exit, steps_h;
        bool on;  // has slots in the current cycle
    };
    std::vector<Run> rs(S);
    for (int i = 0; i < S; ++i) {
        Run &r = rs[i];
        hz_operator *op = ms[i].op;
        r.op = op;
        r.k = ms[i].k;
        r.n = op->n;
        r.nblk = vec_blocks_cols(r.n, r.k);
        r.chunk = vec_chunk(r.n, r.nblk);
        r.V = op->V.as<cx<T>>();
        r.X = op->X.as<cx<T>>();
        r.Bh = op->Bh.as<cx<T>>();
        // Multigrid form: every preconditioned vector the Arnoldi steps
        // compute is kept (slot i of Z), and X holds M y itself, updated from
        // Z at the end of a cycle.  The cycle top and the caller then need no
        // V-cycle of their own on X.
        r.Z = nullptr;
        if constexpr (std::is_same<T, float>::value) {
            if (use_mg) {
                op->Z.ensure((size_t)restart * r.k * r.n * sizeof(cx<float>));
                r.Z = op->Z.as<cx<float>>();
            }
        }
        r.st = gmres_state(op);
        r.member = op->member.as<int>();
        r.dpart = op->dpart.as<double2>();
        r.npart = op->npart.as<double>();
        r.gram = op->gram.as<double>();
        r.early.assign(r.k, -1.0);
        if (!ref_semantics)
            for (int c = 0; c < r.k; ++c) r.early[c] = ms[i].atol[c];
        r.resid.assign(r.k, 0.0);
        r.prev_top.assign(r.k, 0.0);
        r.live_h.assign(r.k, 0);
        r.early_exit.assign(r.k, 0);
        r.steps_h.assign(r.k, 0);
        r.on = true;
    }
    LaunchCache local;
    local.limit = 64;
    LaunchCache &lc = fused ? (batch_cache ? *batch_cache : local) : ms[0].op->graphs;

    {
        Fuser f(s);
        for (int i = 0; i < S; ++i)
            f.add<KZero<T>>(i, dim3(vec_blocks(rs[i].n), rs[i].k), {rs[i].X, rs[i].n, rs[i].n});
    }
    for (Run &r : rs)
        HZ_HIP(hipMemcpyAsync(r.st.tol, r.early.data(), r.k * sizeof(double), hipMemcpyHostToDevice, s));

    // Scenes in the current cycle, as a signature prefix (tag: 0 cycle top,
    // 1 cycle body).
    auto members_sig = [&](int tag) {
        std::vector<uintptr_t> sig = {(uintptr_t)tag, (uintptr_t)S};
        for (int i = 0; i < S; ++i)
            if (rs[i].on) sig.push_back((uintptr_t)i);
        return sig;
    };
    while (true) {
        for (int i = 0; i < S; ++i) {
            Run &r = rs[i];
            if (!r.on) continue;
            for (int c = 0; c < r.k; ++c) r.live_h[c] = ms[i].cs->active[c];
            HZ_HIP(hipMemcpyAsync(r.st.live, r.live_h.data(), r.k * sizeof(int), hipMemcpyHostToDevice, s));
        }
        // Cycle top: V_0 = Bh - R A M y (true residual of the preconditioned
        // system); X is y, or M y itself in the multigrid form.
        {
            std::vector<uintptr_t> sig = members_sig(0);
            for (int i = 0; i < S; ++i)
                if (rs[i].on) {
                    const auto gs = graph_signature(rs[i].op, sizeof(T), mode_op, rs[i].k, restart, false);
                    sig.insert(sig.end(), gs.begin(), gs.end());
                }
            run_sequence(lc, std::move(sig), fused, false, s, [&](Fuser &f) {
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        residual_mode<T>(f, i, r.op, mode_op, r.X, r.Bh, r.V, r.k, r.st.live,
                                         r.npart);
                }
                f.flush();
                // complex64 residual norms over the vector kernels' chunks.
                constexpr bool c64 = std::is_same<T, float>::value;
                if constexpr (c64) {
                    for (int i = 0; i < S; ++i) {
                        Run &r = rs[i];
                        if (r.on)
                            col_norm_partials(f, i, reinterpret_cast<const cx<float> *>(r.V), r.n, r.k,
                                              r.st.live, r.npart, r.nblk, r.chunk);
                    }
                    f.flush();
                }
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        f.add<KCycleStart>(i, dim3(r.k), {r.st, r.npart,
                                                          (int)(c64 ? r.nblk : r.op->tm.nblocks), maxv,
                                                          restart});
                }
                f.flush();
            });
        }
        for (Run &r : rs)
            if (r.on)
                HZ_HIP(hipMemcpyAsync(r.resid.data(), r.st.resid, r.k * sizeof(double), hipMemcpyDeviceToHost, s));
        HZ_HIP(hipStreamSynchronize(s));
        bool any_scene = false;
        for (int i = 0; i < S; ++i) {
            Run &r = rs[i];
            if (!r.on) continue;
            ChunkState &cs = *ms[i].cs;
            const std::vector<double> &atol = ms[i].atol;
            bool any = false;
            for (int c = 0; c < r.k; ++c) {
                if (!cs.active[c]) continue;
                if (verbose())
                    std::fprintf(stderr, "[hz] %s cycle %d scene %d rhs %d: residual %.6e (target %.6e)\n",
                                 sizeof(T) == 8 ? "c128" : "c64", cs.cycles[c], i, c, r.resid[c], atol[c]);
                cs.last_resid[c] = r.resid[c];
                if (!(r.resid[c] > atol[c]) || cs.iterations[c] >= maxiter || !std::isfinite(r.resid[c])) {
                    cs.active[c] = 0;
                    continue;
                }
                // Working-precision floor: the last cycle stopped early on its
                // Givens estimate, yet the true residual did not even halve.
                // More cycles cannot help; the caller's FP64 refinement takes
                // over.
                if (r.early_exit[c] && r.resid[c] > 0.5 * r.prev_top[c]) {
                    cs.active[c] = 0;
                    continue;
                }
                r.prev_top[c] = r.resid[c];
                any = true;
            }
            r.on = any;
            any_scene = any_scene || any;
            if (!any) continue;
            for (int c = 0; c < r.k; ++c) r.live_h[c] = cs.active[c];
            HZ_HIP(hipMemcpyAsync(r.st.live, r.live_h.data(), r.k * sizeof(int), hipMemcpyHostToDevice, s));
            HZ_HIP(hipMemcpyAsync(r.member, r.live_h.data(), r.k * sizeof(int), hipMemcpyHostToDevice, s));
        }
        if (!any_scene) break;
        // The cycle body: `restart` Arnoldi steps, the triangular solve and
        // the iterate update.  Slots that leave early are skipped inside the
        // kernels (device live flags), so the launch sequence is the same
        // every cycle and can be replayed from a captured graph.
        // Two-step blocks (complex64 only): w1 = A M q_j, w2 = A M w1, one dot
        // pass and one update pass over V_0..V_j for both, an intra-block axpy.
        // The V-cycles of basis slot j: Z_j = M V_j.
        auto mg_tasks = [&](int j) {
            std::vector<MgTask> t;
            if constexpr (std::is_same<T, float>::value) {
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (!r.on) continue;
                    t.push_back({i, r.op->mg.get(), r.V + (int64_t)j * r.k * r.n, r.Z + (int64_t)j * r.k * r.n, r.k,
                                 r.st.live, adjoint ? r.op->v32.rs_a : r.op->v32.rs_f});
                }
            }
            return t;
        };
        auto block_step = [&](Fuser &f, int j) {
            if constexpr (std::is_same<T, float>::value) {
                // w1 = A M V_j into slot j+1.
                if (use_mg) {
                    auto t = mg_tasks(j);
                    precondition(f, t, adjoint);
                }
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (!r.on) continue;
                    const cx<float> *in = use_mg ? r.Z + (int64_t)j * r.k * r.n : r.V + (int64_t)j * r.k * r.n;
                    apply_mode<float>(f, i, r.op, mode_op, in, r.n, r.V + (int64_t)(j + 1) * r.k * r.n, r.n, r.k,
                                      r.st.live, r.st.scales + (int64_t)j * r.k);
                }
                f.flush();
                // w2 = A M w1 into slot j+2 (M w1 kept in Z_{j+1}, before w1
                // is orthogonalised).
                if (use_mg) {
                    auto t = mg_tasks(j + 1);
                    precondition(f, t, adjoint);
                }
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (!r.on) continue;
                    const cx<float> *in =
                        use_mg ? r.Z + (int64_t)(j + 1) * r.k * r.n : r.V + (int64_t)(j + 1) * r.k * r.n;
                    apply_mode<float>(f, i, r.op, mode_op, in, r.n, r.V + (int64_t)(j + 2) * r.k * r.n, r.n, r.k,
                                      r.st.live, nullptr);
                }
                f.flush();
                const int groups = (j + 1 + DOT_GROUP / 2 - 1) / (DOT_GROUP / 2);
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (!r.on) continue;
                    const cx<float> *w1 = r.V + (int64_t)(j + 1) * r.k * r.n;
                    const cx<float> *w2 = r.V + (int64_t)(j + 2) * r.k * r.n;
                    const dim3 grid(r.nblk, r.k, groups);
                    typename KDots2<float, 2>::Args a{r.V, r.k, r.n, j + 1, w1, w2, r.st.live,
                                                      r.dpart, maxv, r.chunk,
                                                      r.gram};
                    const int vw = vec_width<float>(r.n, r.k);
                    if (vw == 3)
                        f.add<KDots2<float, 3>>(i, grid, {a.V, a.k, a.n, a.jp1, a.w1, a.w2, a.live, a.partial, a.maxv,
                                                          a.chunk, a.wgram});
                    else if (vw == 2)
                        f.add<KDots2<float, 2>>(i, grid, a);
                    else
                        f.add<KDots2<float, 1>>(i, grid, {a.V, a.k, a.n, a.jp1, a.w1, a.w2, a.live, a.partial, a.maxv,
                                                          a.chunk, a.wgram});
                }
                f.flush();
                for (int i = 0; i < S; ++i)
                    if (rs[i].on)
                        f.add<KReduceDots2>(i, dim3(j + 1, rs[i].k), {rs[i].dpart,
                                                                      (int)rs[i].nblk, maxv, j, rs[i].st, restart});
                f.flush();
                for (int i = 0; i < S; ++i)
                    if (rs[i].on)
                        f.add<KBlockAlpha>(i, dim3(rs[i].k), {rs[i].st, rs[i].gram, (int)rs[i].nblk,
                                                              j, maxv, restart});
                f.flush();
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (!r.on) continue;
                    cx<float> *w1 = r.V + (int64_t)(j + 1) * r.k * r.n;
                    cx<float> *w2 = r.V + (int64_t)(j + 2) * r.k * r.n;
                    const dim3 grid(r.nblk, r.k);
                    typename KUpdate2<float, 2>::Args a{r.V, r.k, r.n, j + 1, w1, w2, r.st.coef, r.st.coef2, maxv,
                                                        r.st.live, r.gram, r.chunk, r.st.alpha};
                    const int vw = vec_width<float>(r.n, r.k);
                    if (vw == 3)
                        f.add<KUpdate2<float, 3>>(i, grid, {a.V, a.k, a.n, a.cnt, a.w1, a.w2, a.coef, a.coef2,
                                                            a.maxv, a.live, a.gram, a.chunk, a.alpha});
                    else if (vw == 2)
                        f.add<KUpdate2<float, 2>>(i, grid, a);
                    else
                        f.add<KUpdate2<float, 1>>(i, grid, {a.V, a.k, a.n, a.cnt, a.w1, a.w2, a.coef, a.coef2, a.maxv,
                                                            a.live, a.gram, a.chunk, a.alpha});
                }
                f.flush();
                for (int i = 0; i < S; ++i)
                    if (rs[i].on)
                        f.add<KBlockFixcoef>(i, dim3(rs[i].k), {rs[i].st, rs[i].gram,
                                                                (int)rs[i].nblk});
                f.flush();
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        f.add<KBlockFixup<float>>(i, dim3(r.nblk, r.k), {r.st, r.n, r.V + (int64_t)(j + 1) * r.k * r.n,
                                                                         r.V + (int64_t)(j + 2) * r.k * r.n,
                                                                         r.gram, r.chunk});
                }
                f.flush();
                for (int i = 0; i < S; ++i)
                    if (rs[i].on)
                        f.add<KBlockFinish>(i, dim3(rs[i].k), {rs[i].st, rs[i].gram,
                                                               (int)rs[i].nblk, j, restart});
                f.flush();
            }
        };
        auto body = [&](Fuser &f) {
            int j0 = 0;
            if (two)
                for (; j0 + 1 < restart; j0 += 2) block_step(f, j0);
            for (int j = j0; j < restart; ++j) {
                if constexpr (std::is_same<T, float>::value) {
                    if (use_mg) {
                        auto t = mg_tasks(j);
                        precondition(f, t, adjoint);
                    }
                }
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (!r.on) continue;
                    const cx<T> *in = r.V + (int64_t)j * r.k * r.n;
                    if constexpr (std::is_same<T, float>::value)
                        if (use_mg) in = r.Z + (int64_t)j * r.k * r.n;
                    apply_mode<T>(f, i, r.op, mode_op, in, r.n, r.V + (int64_t)(j + 1) * r.k * r.n, r.n, r.k,
                                  r.st.live, r.st.scales + (int64_t)j * r.k);
                }
                f.flush();
                const int groups = (j + 1 + dot_group<T>() - 1) / dot_group<T>();
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        launch_dots<T>(f, i, dim3(r.nblk, r.k, groups), r.V, r.k, r.n, j + 1,
                                       r.V + (int64_t)(j + 1) * r.k * r.n, r.st.live, r.dpart, maxv,
                                       r.chunk);
                }
                f.flush();
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        f.add<KReduceDots>(i, dim3(j + 1, r.k), {r.dpart, (int)r.nblk, maxv, j,
                                                                 r.st.scales, r.st.live, r.st.H, r.st.coef, restart});
                }
                f.flush();
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        launch_update<T>(f, i, dim3(r.nblk, r.k), r.V, r.k, r.n, j + 1, nullptr,
                                         r.V + (int64_t)(j + 1) * r.k * r.n, r.n, r.st.coef, maxv, r.st.live,
                                         r.npart, r.chunk);
                }
                f.flush();
                for (int i = 0; i < S; ++i) {
                    Run &r = rs[i];
                    if (r.on)
                        f.add<KArnoldiFinish>(i, dim3(r.k), {r.st, r.npart, (int)r.nblk, j, maxv,
                                                             restart});
                }
                f.flush();
            }
            for (int i = 0; i < S; ++i)
                if (rs[i].on)
                    f.add<KTrisolve>(i, dim3(rs[i].k), {rs[i].st, maxv, restart, rs[i].member, (int)(use_mg && two)});
            f.flush();
            for (int i = 0; i < S; ++i) {
                Run &r = rs[i];
                if (!r.on) continue;
                const cx<T> *basis = r.V;
                if constexpr (std::is_same<T, float>::value)
                    if (use_mg) basis = r.Z;
                launch_update<T>(f, i, dim3(r.nblk, r.k), basis, r.k, r.n, 0, r.st.steps, r.X, r.n, r.st.coef, maxv,
                                 r.member, nullptr, r.chunk);
            }
            f.flush();
        };
        {
            std::vector<uintptr_t> sig = members_sig(1);
            for (int i = 0; i < S; ++i)
                if (rs[i].on) {
                    const auto gs = graph_signature(rs[i].op, sizeof(T), mode_op + 16 * (two ? 1 : 0), rs[i].k,
                                                    restart, use_mg);
                    sig.insert(sig.end(), gs.begin(), gs.end());
                }
            run_sequence(lc, std::move(sig), fused, use_graph, s, body);
        }
        for (Run &r : rs)
            if (r.on)
                HZ_HIP(hipMemcpyAsync(r.steps_h.data(), r.st.steps, r.k * sizeof(int), hipMemcpyDeviceToHost, s));
        HZ_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < S; ++i) {
            Run &r = rs[i];
            if (!r.on) continue;
            ChunkState &cs = *ms[i].cs;
            for (int c = 0; c < r.k; ++c) {
                if (!cs.active[c]) continue;
                cs.iterations[c] += ref_semantics ? restart : r.steps_h[c];
                cs.cycles[c] += 1;
                r.early_exit[c] = !ref_semantics && r.steps_h[c] < restart;
            }
        }
    }
}

void reserve_workspace(hz_operator *op, int k, int restart, size_t elem) {
    const int64_t n = op->n;
    const int maxv = restart + 1;
    const int64_t nblk = vec_blocks_cols(n, k);
    op->V.ensure((size_t)maxv * k * n * elem);
    op->X.ensure((size_t)k * n * elem);
    op->Bh.ensure((size_t)k * n * elem);
    op->dpart.ensure((size_t)2 * k * maxv * nblk * sizeof(double2));  // two targets in a block step
    op->npart.ensure((size_t)k * std::max<int64_t>(nblk, op->tm.nblocks) * 2 * sizeof(double));
    op->cpart.ensure((size_t)k * op->tm.nblocks * 6 * sizeof(double));
    op->red.ensure((size_t)k * 8 * sizeof(double));
    op->Hm.ensure((size_t)k * (restart + 1) * restart * sizeof(cx<double>));
    op->gc.ensure((size_t)k * restart * sizeof(double));
    op->gs.ensure((size_t)k * restart * sizeof(cx<double>));
    op->g.ensure((size_t)k * (restart + 1) * sizeof(cx<double>));
    op->scales.ensure((size_t)maxv * k * sizeof(double));
    op->coef.ensure((size_t)k * maxv * sizeof(cx<double>));
    op->coef2.ensure((size_t)k * maxv * sizeof(cx<double>));
    op->Hu.ensure((size_t)k * (restart + 1) * restart * sizeof(cx<double>));
    op->alpha.ensure((size_t)2 * k * sizeof(cx<double>));
    op->bfix.ensure((size_t)2 * k * sizeof(int));
    op->gram.ensure((size_t)k * nblk * 4 * sizeof(double));
    op->resid.ensure((size_t)k * sizeof(double));
    op->tol.ensure((size_t)k * sizeof(double));
    op->steps.ensure((size_t)k * sizeof(int));
    op->live.ensure((size_t)k * sizeof(int));
    op->member.ensure((size_t)k * sizeof(int));
}

struct CheckResult {
    double se, sb, ue, ub, bwd, bad;
};

// FP64 acceptance check of x (physical) against b for k columns; optionally
// writes the scaled refinement residual into op->rhs32.  Launches only: the
// sums land in op->red (read by check_read after a stream synchronisation).
void check_launch(Fuser &f, int sc, hz_operator *op, int adjoint, const cx<double> *x, int64_t ldx,
                  const cx<double> *b, int64_t ldb, int k, cx<float> *rhs_out) {
    const dim3 grid(op->tm.nblocks, k);
    double *part = op->cpart.as<double>();
    if (adjoint)
        f.add<KCheck<MODE_AH>>(sc, grid, {op->v64, op->tm, x, ldx, b, ldb, nullptr, part, rhs_out, op->n});
    else
        f.add<KCheck<MODE_A>>(sc, grid, {op->v64, op->tm, x, ldx, b, ldb, nullptr, part, rhs_out, op->n});
}

void check_reduce(Fuser &f, int sc, hz_operator *op, int k) {
    f.add<KReduceRows>(sc, dim3(6, k), {op->cpart.as<double>(), (int)op->tm.nblocks, 6, 1 << 4, op->red.as<double>()});
}

std::vector<CheckResult> check_read(const std::vector<double> &h, int k) {
    std::vector<CheckResult> out(k);
    for (int r = 0; r < k; ++r) {
        const double *v = h.data() + (size_t)r * 6;
        out[r] = {v[0], v[1], v[2], v[3], v[4], v[5]};
    }
    return out;
}

std::vector<CheckResult> check(hz_operator *op, int adjoint, const cx<double> *x, int64_t ldx,
                               const cx<double> *b, int64_t ldb, int k, cx<float> *rhs_out, hipStream_t s) {
    Fuser f(s);
    check_launch(f, 0, op, adjoint, x, ldx, b, ldb, k, rhs_out);
    check_reduce(f, 0, op, k);
    std::vector<double> h((size_t)k * 6);
    HZ_HIP(hipMemcpyAsync(h.data(), op->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HZ_HIP(hipStreamSynchronize(s));
    return check_read(h, k);
}

void fill_stats(hz_gmres_stats &st, const CheckResult &c, const hz_gmres_opts &o) {
    if (verbose())
        std::fprintf(stderr, "[hz] check: |Re|^2 %.6e |Rb|^2 %.6e |e|^2 %.6e |b|^2 %.6e bwd %.6e nonfinite %g\n", c.se,
                     c.sb, c.ue, c.ub, c.bwd, c.bad);
    st.scaled_residual = c.sb > 0 ? std::sqrt(c.se / c.sb) : 0.0;
    st.relative_residual = c.ub > 0 ? std::sqrt(c.ue / c.ub) : 0.0;
    st.backward_error = c.bwd;
    st.finite = c.bad == 0 ? 1 : 0;
    st.converged = (st.info == 0 && st.finite && st.scaled_residual <= o.rtol &&
                    st.backward_error <= o.backward_error_tol &&
                    !(o.relative_rtol > 0 && st.relative_residual > o.relative_rtol))
                       ? 1
                       : 0;
}

// Plain form: FP64 GMRES with the reference's semantics on one chunk.
void solve_plain(hz_operator *op, const hz_gmres_opts &o, const cx<double> *b, int64_t ldb, cx<double> *x,
                 int64_t ldx, int k, hz_gmres_stats *stats, hipStream_t s) {
    const int64_t n = op->n;
    reserve_workspace(op, k, o.restart, sizeof(cx<double>));
    const double *rs = o.adjoint ? op->rsa64.as<double>() : op->rsf64.as<double>();
    const cx<double> *m = o.adjoint ? op->ma64.as<cx<double>>() : op->mf64.as<cx<double>>();
    const int64_t nblk = vec_blocks_cols(n, k);
    const int64_t chunk = vec_chunk(n, nblk);
    double *part = op->npart.as<double>();
    {
        Fuser f(s);
        f.add<KScaleRhs<double>>(0, dim3(nblk, k), {b, ldb, rs, op->Bh.as<cx<double>>(), n, n, part, chunk});
        f.add<KReduceRows>(0, dim3(2, k), {part, (int)nblk, 2, 0, op->red.as<double>()});
    }
    std::vector<double> norms((size_t)2 * k);
    HZ_HIP(hipMemcpyAsync(norms.data(), op->red.p, norms.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HZ_HIP(hipStreamSynchronize(s));
    ChunkState cs;
    cs.k = k;
    cs.active.assign(k, 1);
    cs.iterations.assign(k, 0);
    cs.cycles.assign(k, 0);
    cs.last_resid.assign(k, 0.0);
    std::vector<double> atol(k);
    std::vector<int> zero_rhs(k, 0);
    for (int r = 0; r < k; ++r) {
        const double bn = std::sqrt(norms[r * 2 + 1]);
        atol[r] = o.rtol * std::sqrt(norms[r * 2]);
        if (bn == 0.0) {
            zero_rhs[r] = 1;
            cs.active[r] = 0;
        }
    }
    std::vector<Member> ms(1);
    ms[0].op = op;
    ms[0].k = k;
    ms[0].cs = &cs;
    ms[0].atol = atol;
    gmres_cycles<double>(ms, o.adjoint ? MODE_RAMH : MODE_RAM, o.restart, o.maxiter, true, s, PRE_JACOBI,
                         o.use_graph != 0);
    // x = M y for every column (zero right-hand sides give y = 0).
    {
        Fuser f(s);
        f.add<KPrecondOut<double>>(0, dim3(vec_blocks(n), k), {(const cx<double> *)op->X.as<cx<double>>(), n, m, x,
                                                               ldx, n, nullptr, 0});
    }
    const auto chk = check(op, o.adjoint, x, ldx, b, ldb, k, nullptr, s);
    for (int r = 0; r < k; ++r) {
        hz_gmres_stats &st = stats[r];
        std::memset(&st, 0, sizeof(st));
        if (zero_rhs[r]) {
            st.finite = 1;
            st.converged = 1;
            continue;
        }
        st.iterations = cs.iterations[r];
        st.cycles = cs.cycles[r];
        const bool conv = !(cs.last_resid[r] > atol[r]);
        st.info = (cs.iterations[r] == o.maxiter && !conv) ? cs.iterations[r] : 0;
        if (!std::isfinite(cs.last_resid[r])) st.info = std::max(1, cs.iterations[r]);
        fill_stats(st, chk[r], o);
    }
}

// One scene's chunk of a mixed-form solve: FP64 right-hand sides and
// solution (caller's device buffers), its stats and its refinement progress.
struct MixedJob {
    hz_operator *op = nullptr;
    const cx<double> *b = nullptr;
    cx<double> *x = nullptr;
    int64_t ldb = 0, ldx = 0;
    int k = 0;
    hz_gmres_stats *stats = nullptr;
    ChunkState cs;
    std::vector<int> done, refinements, zero_rhs, member;
    std::vector<CheckResult> last;
};

// Mixed form: complex64 GMRES inner solves under FP64 iterative refinement,
// for one chunk of each job's scene.  Several jobs refine in lockstep: each
// round checks every scene, then runs one lockstep inner solve over the
// columns that still need one (gmres_cycles with several members).
void solve_mixed(std::vector<MixedJob> &jobs, const hz_gmres_opts &o, hipStream_t s, LaunchCache *batch_cache) {
    const int S = (int)jobs.size();
    const bool fused = S > 1;
    const bool use_mg = o.preconditioner == PRE_MULTIGRID;
    {
        Fuser f(s);
        for (int i = 0; i < S; ++i) {
            MixedJob &jb = jobs[i];
            hz_operator *op = jb.op;
            reserve_workspace(op, jb.k, o.restart, sizeof(cx<float>));
            op->rhs32.ensure((size_t)jb.k * op->n * sizeof(cx<float>));
            if (use_mg) mg_reserve(*op->mg, jb.k);
            f.add<KZero<double>>(i, dim3(vec_blocks(op->n), jb.k), {jb.x, jb.ldx, op->n});
            // Per-RHS count of two-step blocks that took the measured coefficient.
            const std::vector<int> zeros((size_t)2 * jb.k, 0);
            HZ_HIP(hipMemcpyAsync(op->bfix.p, zeros.data(), zeros.size() * sizeof(int), hipMemcpyHostToDevice, s));
            HZ_HIP(hipStreamSynchronize(s));
            jb.cs.k = jb.k;
            jb.cs.active.assign(jb.k, 1);
            jb.cs.iterations.assign(jb.k, 0);
            jb.cs.cycles.assign(jb.k, 0);
            jb.cs.last_resid.assign(jb.k, 0.0);
            jb.done.assign(jb.k, 0);
            jb.refinements.assign(jb.k, 0);
            jb.zero_rhs.assign(jb.k, 0);
            jb.member.assign(jb.k, 0);
        }
    }
    std::vector<std::vector<double>> sums(S);
    for (int ref = 0;; ++ref) {
        // FP64 check of every scene (refinement residual into rhs32).
        {
            Fuser f(s);
            for (int i = 0; i < S; ++i)
                check_launch(f, i, jobs[i].op, o.adjoint, jobs[i].x, jobs[i].ldx, jobs[i].b, jobs[i].ldb, jobs[i].k,
                             jobs[i].op->rhs32.as<cx<float>>());
            for (int i = 0; i < S; ++i) check_reduce(f, i, jobs[i].op, jobs[i].k);
            for (int i = 0; i < S; ++i) {
                sums[i].resize((size_t)jobs[i].k * 6);
                HZ_HIP(hipMemcpyAsync(sums[i].data(), jobs[i].op->red.p, sums[i].size() * sizeof(double),
                                      hipMemcpyDeviceToHost, s));
            }
            HZ_HIP(hipStreamSynchronize(s));
        }
        std::vector<Member> ms;
        std::vector<int> job_of;
        for (int i = 0; i < S; ++i) {
            MixedJob &jb = jobs[i];
            jb.last = check_read(sums[i], jb.k);
            std::vector<double> atol(jb.k, 0.0);
            bool any = false;
            for (int r = 0; r < jb.k; ++r) {
                if (jb.done[r]) continue;
                const CheckResult &c = jb.last[r];
                if (c.ub == 0.0) {
                    jb.zero_rhs[r] = 1;
                    jb.done[r] = 1;
                    continue;
                }
                const double scaled = c.sb > 0 ? std::sqrt(c.se / c.sb) : 0.0;
                const double rel = std::sqrt(c.ue / c.ub);
                const bool rel_ok = !(o.relative_rtol > 0 && rel > o.relative_rtol);
                const bool ok = c.bad == 0 && scaled <= o.rtol && c.bwd <= o.backward_error_tol && rel_ok;
                if (verbose())
                    std::fprintf(stderr,
                                 "[hz] refinement %d scene %d rhs %d: scaled residual %.3e relative residual %.3e "
                                 "backward error %.3e%s\n",
                                 ref, i, r, scaled, rel, c.bwd, ok ? " (accepted)" : "");
                if (ok || ref >= o.max_refinements || jb.cs.iterations[r] >= o.maxiter || c.bad != 0) {
                    jb.done[r] = 1;
                    continue;
                }
                // Aim below the scaled-residual target.  When only the
                // componentwise backward error fails, ask for the reduction it
                // needs with a factor of 2 to spare (at least 2x, at most
                // 100x): it typically misses by less than 3x, and every decade
                // asked for costs a fresh Krylov space dozens of steps.
                // A relative-residual miss (when asked for) likewise asks for
                // the reduction it needs, in the scaled norm the inner solve
                // measures.
                auto reduce = [&](double need) {
                    return std::sqrt(c.se) * std::min(0.5, std::max(0.01, 0.5 * need));
                };
                double goal = scaled > o.rtol                   ? 0.5 * o.rtol * std::sqrt(c.sb)
                              : c.bwd > o.backward_error_tol ? reduce(o.backward_error_tol / c.bwd)
                                                                : HUGE_VAL;
                if (!rel_ok) goal = std::min(goal, reduce(o.relative_rtol / rel));
                atol[r] = std::max(goal, o.inner_rtol * std::sqrt(c.se));
                any = true;
            }
            if (!any) continue;
            // The inner right-hand side is the scaled residual written by the check.
            HZ_HIP(hipMemcpyAsync(jb.op->Bh.p, jb.op->rhs32.p, (size_t)jb.k * jb.op->n * sizeof(cx<float>),
                                  hipMemcpyDeviceToDevice, s));
            for (int r = 0; r < jb.k; ++r) jb.cs.active[r] = jb.member[r] = jb.done[r] ? 0 : 1;
            Member m;
            m.op = jb.op;
            m.k = jb.k;
            m.cs = &jb.cs;  // the iteration budget is cumulative over refinements
            m.atol = atol;
            ms.push_back(std::move(m));
            job_of.push_back(i);
        }
        if (ms.empty()) break;
        gmres_cycles<float>(ms, o.adjoint ? MODE_RAMH : MODE_RAM, o.restart, o.maxiter, false, s, o.preconditioner,
                            o.use_graph != 0, o.arnoldi_block, fused ? batch_cache : nullptr);
        Fuser f(s);
        for (size_t q = 0; q < ms.size(); ++q) {
            MixedJob &jb = jobs[job_of[q]];
            for (int r = 0; r < jb.k; ++r)
                if (jb.member[r]) jb.refinements[r] += 1;
            hz_operator *op = jb.op;
            HZ_HIP(hipMemcpyAsync(op->member.p, jb.member.data(), jb.k * sizeof(int), hipMemcpyHostToDevice, s));
            if (use_mg) {
                // x += M y: X already holds it.
                f.add<KAccumulate>(job_of[q], dim3(vec_blocks(op->n), jb.k),
                                   {(const cx<float> *)op->X.as<cx<float>>(), op->n, jb.x, jb.ldx,
                                    (const int *)op->member.as<int>()});
            } else {
                const cx<double> *m = o.adjoint ? op->ma64.as<cx<double>>() : op->mf64.as<cx<double>>();
                f.add<KPrecondOut<float>>(job_of[q], dim3(vec_blocks(op->n), jb.k),
                                          {(const cx<float> *)op->X.as<cx<float>>(), op->n, m, jb.x, jb.ldx, op->n,
                                           (const int *)op->member.as<int>(), 1});
            }
        }
    }
    for (int i = 0; i < S; ++i) {
        MixedJob &jb = jobs[i];
        if (verbose() && o.arnoldi_block == 2) {
            std::vector<int> fixes((size_t)2 * jb.k);
            HZ_HIP(hipMemcpyAsync(fixes.data(), jb.op->bfix.p, fixes.size() * sizeof(int), hipMemcpyDeviceToHost, s));
            HZ_HIP(hipStreamSynchronize(s));
            for (int r = 0; r < jb.k; ++r)
                if (fixes[jb.k + r])
                    std::fprintf(stderr, "[hz] scene %d rhs %d: %d two-step blocks took the measured coefficient\n", i,
                                 r, fixes[jb.k + r]);
        }
        for (int r = 0; r < jb.k; ++r) {
            hz_gmres_stats &st = jb.stats[r];
            std::memset(&st, 0, sizeof(st));
            if (jb.zero_rhs[r]) {
                st.finite = 1;
                st.converged = 1;
                continue;
            }
            st.iterations = jb.cs.iterations[r];
            st.cycles = jb.cs.cycles[r];
            st.refinements = jb.refinements[r];
            fill_stats(st, jb.last[r], o);
            if (!st.converged) {
                st.info = std::max(1, jb.cs.iterations[r]);
                st.converged = 0;
            }
        }
    }
}

void check_opts(const hz_operator *op, const hz_gmres_opts &o) {
    if (o.restart < 1 || o.restart > 63) throw BadArgument("restart must be in [1, 63]");
    if (o.maxiter < 0) throw BadArgument("maxiter must be >= 0");
    if (!(o.rtol > 0) || !(o.backward_error_tol > 0)) throw BadArgument("tolerances must be positive");
    if (!(o.relative_rtol >= 0)) throw BadArgument("relative_rtol must be >= 0");
    if (o.precision != HZ_C128 && o.precision != HZ_C64) throw BadArgument("unknown precision");
    if (o.arnoldi_block != 1 && o.arnoldi_block != 2) throw BadArgument("arnoldi_block must be 1 or 2");
    if (o.arnoldi_block == 2 && o.precision != HZ_C64) throw BadArgument("two-step blocks run in the mixed form");
    if (o.preconditioner == PRE_MULTIGRID) {
        if (o.precision != HZ_C64) throw BadArgument("the multigrid preconditioner runs in the mixed form only");
        if (!op->mg) throw BadArgument("no multigrid hierarchy set (hz_set_multigrid)");
    } else if (o.preconditioner != PRE_JACOBI) {
        throw BadArgument("unknown preconditioner");
    }
}

// Right-hand sides per batch; the stencil kernels stage at most APPLY_MAX_RHS in LDS.
int chunk_width(const hz_gmres_opts &o, int nrhs) {
    return std::max(1, std::min<int>(std::min<int>(nrhs, APPLY_MAX_RHS), o.chunk > 0 ? o.chunk : 1));
}

// Validated hz_gmres on one stream (throws on bad arguments or HIP errors).
void gmres_run(hz_operator *op, const hz_gmres_opts *opts, const void *b, void *x, int nrhs, int64_t ld,
               hz_gmres_stats *stats, hipStream_t s) {
    if (!op) throw BadArgument("null operator handle");
    if (!opts || !b || !x || !stats) throw BadArgument("opts, b, x and stats are required");
    const hz_gmres_opts &o = *opts;
    if (nrhs < 1 || ld < op->n) throw BadArgument("need nrhs >= 1 and ld >= N");
    check_opts(op, o);
    const int chunk = chunk_width(o, nrhs);
    for (int c0 = 0; c0 < nrhs; c0 += chunk) {
        const int k = std::min(chunk, nrhs - c0);
        const cx<double> *bc = static_cast<const cx<double> *>(b) + (int64_t)c0 * ld;
        cx<double> *xc = static_cast<cx<double> *>(x) + (int64_t)c0 * ld;
        if (o.precision == HZ_C128) {
            solve_plain(op, o, bc, ld, xc, ld, k, stats + c0, s);
        } else {
            std::vector<MixedJob> jobs(1);
            jobs[0].op = op;
            jobs[0].b = bc;
            jobs[0].x = xc;
            jobs[0].ldb = jobs[0].ldx = ld;
            jobs[0].k = k;
            jobs[0].stats = stats + c0;
            solve_mixed(jobs, o, s, nullptr);
        }
    }
}

// ---------------------------------------------------------------- autotuner

// Mean duration (ms) of `reps` back-to-back launches after one warm-up.
template <typename F>
float time_launches(F &&launch, int reps, hipStream_t s) {
    hipEvent_t e0, e1;
    HZ_HIP(hipEventCreate(&e0));
    HZ_HIP(hipEventCreate(&e1));
    launch();
    HZ_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) launch();
    HZ_HIP(hipEventRecord(e1, s));
    HZ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    HZ_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

// Keeps the device busy with `launch` for about `seconds`: the GPU's clocks
// ramp up under sustained load and drop within a second of idling, and the
// timings that follow should all see the clocks a running solve sees.
template <typename F>
void warm_up(F &&launch, double seconds, hipStream_t s) {
    const float ms = time_launches(launch, 3, s);
    const int reps = (int)std::min(20000.0, std::max(1.0, seconds * 1e3 / std::max(ms, 1e-3f)));
    for (int i = 0; i < reps; ++i) launch();
    HZ_HIP(hipStreamSynchronize(s));
}

// Fastest slab shape for one kernel: `launch(shape)` runs it once.  A shape
// replaces the current choice only when it is more than 3 % faster, so that
// timing noise does not flip between equivalent shapes (earlier shapes in the
// table win ties; `preferred` is tried first).
template <typename F>
int pick_shape(F &&launch, hipStream_t s, int preferred = 0) {
    int best = preferred;
    float best_ms = time_launches([&] { launch(preferred); }, 5, s);
    for (int sh = 0; sh < SLAB_SHAPES; ++sh) {
        const float ms = sh == preferred ? best_ms : time_launches([&] { launch(sh); }, 5, s);
        if (verbose())
            std::fprintf(stderr, "[hz] autotune shape %d (zb %d, ty %d, rp %d, nt %d, flat %d): %.1f us\n", sh,
                         SLAB_SHAPE[sh].zb, SLAB_SHAPE[sh].ty, SLAB_SHAPE[sh].rp, SLAB_SHAPE[sh].nt,
                         SLAB_SHAPE[sh].flat, ms * 1e3);
        if (ms < 0.97f * best_ms) {
            best = sh;
            best_ms = ms;
        }
    }
    return best;
}

// Device bytes a mixed-form chunk of k columns needs besides the caller's
// FP64 x and b: the complex64 basis, iterate, right-hand sides and the
// multigrid level vectors.
size_t chunk_bytes(const hz_operator *op, int k, int restart) {
    size_t nodes = (size_t)op->n * (restart + 1 + 3);
    if (op->mg) {
        nodes += (size_t)op->n * restart;  // preconditioned basis slots
        for (const auto &L : op->mg->levels) nodes += 3 * (size_t)L->n;
    }
    return nodes * k * sizeof(cx<float>);
}

void drop_graphs(hz_operator *op) { op->graphs.clear(); }

void release_workspace(hz_operator *op) {
    for (DevBuf *b : {&op->V, &op->X, &op->Z, &op->Bh, &op->rhs32}) b->release();
    if (op->mg) {
        for (auto &L : op->mg->levels) {
            L->bb.release();
            L->xx.release();
            L->tt.release();
        }
        op->mg->kcap = 0;
    }
}

// One Arnoldi step of the mixed form on k columns at basis index j (timing
// only: preconditioner, stencil, Gram-Schmidt dots and update).
void arnoldi_step_for_timing(hz_operator *op, int k, int j, int restart, bool use_mg, hipStream_t s) {
    const int64_t n = op->n;
    const int maxv = restart + 1;
    const int64_t nblk = vec_blocks_cols(n, k);
    const int64_t chunk = vec_chunk(n, nblk);
    cx<float> *V = op->V.as<cx<float>>();
    const cx<float> *in = V + (int64_t)j * k * n;
    cx<float> *w = V + (int64_t)(j + 1) * k * n;
    const int *live = op->live.as<int>();
    Fuser f(s);
    if (use_mg) {
        cx<float> *z = op->mg->levels[0]->xx.as<cx<float>>();
        std::vector<MgTask> t = {{0, op->mg.get(), in, z, k, live, op->v32.rs_f}};
        precondition(f, t, false);
        in = z;
    }
    apply_mode<float>(f, 0, op, use_mg ? MODE_RA : MODE_RAM, in, n, w, n, k, live, op->scales.as<double>());
    const int groups = (j + 1 + DOT_GROUP - 1) / DOT_GROUP;
    launch_dots<float>(f, 0, dim3(nblk, k, groups), V, k, n, j + 1, w, live, op->dpart.as<double2>(), maxv, chunk);
    launch_update<float>(f, 0, dim3(nblk, k), V, k, n, j + 1, nullptr, w, n, op->coef.as<cx<double>>(), maxv, live,
                         op->npart.as<double>(), chunk);
}

template <typename F>
hz_status guarded(F &&fn) {
    try {
        fn();
        return HZ_OK;
    } catch (const BadArgument &e) {
        g_error = e.what();
        return HZ_BAD_ARGUMENT;
    } catch (const HipFailure &e) {
        g_error = e.what();
        return e.code == hipErrorOutOfMemory ? HZ_OUT_OF_MEMORY : HZ_DEVICE_ERROR;
    } catch (const std::bad_alloc &) {
        g_error = "host allocation failed";
        return HZ_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_error = e.what();
        return HZ_DEVICE_ERROR;
    }
}

void check_op(const hz_operator *op) {
    if (!op) throw BadArgument("null operator handle");
}

// Partial-sum workspace of the response-basis kernels, one per device,
// grow-only for the life of the process.
DevBuf &basis_workspace(size_t bytes) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<DevBuf>> bufs;
    int dev = 0;
    HZ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto &b = bufs[dev];
    if (!b) b = std::make_unique<DevBuf>();
    b->ensure(bytes);
    return *b;
}

void check_basis_args(const void *G, int64_t ldg, int64_t nrows, int32_t m) {
    if (!G) throw BadArgument("G must be a device pointer");
    if (nrows < 1 || m < 1 || ldg < m) throw BadArgument("need nrows >= 1, m >= 1 and ldg >= m");
    if (m > BASIS_MAX_M) throw BadArgument("at most 2048 emitters per basis row");
}

}  // namespace

extern "C" {

const char *hz_last_error(void) { return g_error.c_str(); }

hz_status hz_create(const hz_grid_desc *d, hz_operator **out) {
    if (!out) {
        g_error = "out is null";
        return HZ_BAD_ARGUMENT;
    }
    *out = nullptr;
    hz_operator *op = new (std::nothrow) hz_operator();
    if (!op) {
        g_error = "host allocation failed";
        return HZ_OUT_OF_MEMORY;
    }
    const hz_status st = guarded([&] {
        if (!d) throw BadArgument("descriptor is null");
        if (d->nx < 2 || d->ny < 2 || d->nz < 2) throw BadArgument("every grid edge needs >= 2 nodes");
        if (!(d->dx > 0) || !(d->omega > 0) || !std::isfinite(d->dx) || !std::isfinite(d->omega))
            throw BadArgument("dx and omega must be finite and positive");
        if (!d->rho || !d->kappa) throw BadArgument("rho and kappa are required");
        op->nx = d->nx;
        op->ny = d->ny;
        op->nz = d->nz;
        op->n = (int64_t)d->nx * d->ny * d->nz;
        op->dx = d->dx;
        op->omega = d->omega;
        const int dims[3] = {d->nx, d->ny, d->nz};
        for (int f = 0; f < 6; ++f) {
            if (d->face_kind[f] < HZ_FACE_REFLECTING || d->face_kind[f] > HZ_FACE_IMPEDANCE)
                throw BadArgument("unknown face kind");
            if (d->face_kind[f] == HZ_FACE_SOFT) op->soft |= 1 << f;
            if (d->face_kind[f] == HZ_FACE_OPEN) op->open_mask |= 1 << f;
            const bool needs = d->face_kind[f] == HZ_FACE_OPEN || d->face_kind[f] == HZ_FACE_IMPEDANCE;
            if (needs && !d->face_term[f]) throw BadArgument("open/impedance face needs a face-term table");
        }
        for (int64_t p = 0; p < op->n; ++p)
            if (!(d->rho[p] > 0) || !std::isfinite(d->rho[p]) || !(d->kappa[p] > 0) || !std::isfinite(d->kappa[p]))
                throw BadArgument("rho and kappa must be finite and positive");
        std::vector<float> f32((size_t)op->n);
        upload(op->rho64, d->rho, op->n);
        upload(op->kap64, d->kappa, op->n);
        for (int64_t p = 0; p < op->n; ++p) f32[p] = (float)d->rho[p];
        upload(op->rho32, f32.data(), op->n);
        for (int64_t p = 0; p < op->n; ++p) f32[p] = (float)d->kappa[p];
        upload(op->kap32, f32.data(), op->n);
        const double h2 = d->dx * d->dx;
        for (int ax = 0; ax < 3; ++ax) {
            const int na = dims[ax];
            if (!d->weight[ax] || !d->stretch[ax]) throw BadArgument("per-axis weight and stretch are required");
            std::vector<cx<double>> a(na), q(std::max(na - 1, 1));
            std::vector<cx<float>> a32(na), q32(std::max(na - 1, 1));
            for (int i = 0; i < na; ++i) {
                const cx<double> sv = mk<double>(d->stretch[ax][2 * i], d->stretch[ax][2 * i + 1]);
                a[i] = d->weight[ax][i] * sv;
                a32[i] = cast<float>(a[i]);
            }
            for (int i = 0; i + 1 < na; ++i) {
                const cx<double> s0 = mk<double>(d->stretch[ax][2 * i], d->stretch[ax][2 * i + 1]);
                const cx<double> s1 = mk<double>(d->stretch[ax][2 * i + 2], d->stretch[ax][2 * i + 3]);
                // 2 / (mid * h^2) with mid = (s_i + s_{i+1}) / 2
                q[i] = crecip((0.5 * h2) * (s0 + s1));
                q[i] = 2.0 * q[i];
                q32[i] = cast<float>(q[i]);
            }
            upload(op->a64[ax], a.data(), na);
            upload(op->q64[ax], q.data(), q.size());
            upload(op->a32[ax], a32.data(), na);
            upload(op->q32[ax], q32.data(), q32.size());
        }
        for (int f = 0; f < 6; ++f) {
            if (!(d->face_kind[f] == HZ_FACE_OPEN || d->face_kind[f] == HZ_FACE_IMPEDANCE)) continue;
            const int ax = f / 2;
            const int b0 = ax == 0 ? 1 : 0, c0 = ax == 2 ? 1 : 2;
            const int64_t cnt = (int64_t)dims[b0] * dims[c0];
            upload(op->face64[f], reinterpret_cast<const cx<double> *>(d->face_term[f]), cnt);
            std::vector<cx<float>> t32(cnt);
            for (int64_t e = 0; e < cnt; ++e)
                t32[e] = mk<float>((float)d->face_term[f][2 * e], (float)d->face_term[f][2 * e + 1]);
            upload(op->face32[f], t32.data(), cnt);
        }
        for (DevBuf *bf : {&op->rsf64, &op->rsa64}) bf->ensure(op->n * sizeof(double));
        for (DevBuf *bf : {&op->mf64, &op->ma64}) bf->ensure(op->n * sizeof(cx<double>));
        for (DevBuf *bf : {&op->rsf32, &op->rsa32}) bf->ensure(op->n * sizeof(float));
        for (DevBuf *bf : {&op->mf32, &op->ma32}) bf->ensure(op->n * sizeof(cx<float>));

        auto fill = [&](auto &v, auto *rho, auto *kap, DevBuf *a, DevBuf *q, DevBuf *face, DevBuf &rsf, DevBuf &mf,
                        DevBuf &rsa, DevBuf &ma) {
            using TT = std::remove_pointer_t<decltype(rho)>;
            v.nx = op->nx;
            v.ny = op->ny;
            v.nz = op->nz;
            v.n = op->n;
            v.omega2 = (TT)(op->omega * op->omega);
            v.omega2_im = 0;
            v.rho = rho;
            v.kappa = kap;
            for (int ax = 0; ax < 3; ++ax) {
                v.a[ax] = a[ax].template as<cx<TT>>();
                v.q[ax] = q[ax].template as<cx<TT>>();
            }
            for (int f = 0; f < 6; ++f) v.face[f] = face[f].p ? face[f].template as<cx<TT>>() : nullptr;
            v.soft = op->soft;
            v.rs_f = rsf.template as<TT>();
            v.m_f = mf.template as<cx<TT>>();
            v.rs_a = rsa.template as<TT>();
            v.m_a = ma.template as<cx<TT>>();
        };
        fill(op->v64, op->rho64.as<double>(), op->kap64.as<double>(), op->a64, op->q64, op->face64, op->rsf64,
             op->mf64, op->rsa64, op->ma64);
        fill(op->v32, op->rho32.as<float>(), op->kap32.as<float>(), op->a32, op->q32, op->face32, op->rsf32,
             op->mf32, op->rsa32, op->ma32);
        op->tm = make_tilemap(op->nx, op->ny, op->nz, 4);
        const int64_t blocks = (op->n + 255) / 256;
        hipLaunchKernelGGL(k_setup, dim3(blocks), dim3(256), 0, 0, op->v64, op->rsf64.as<double>(),
                           op->mf64.as<cx<double>>(), op->rsa64.as<double>(), op->ma64.as<cx<double>>(),
                           op->rsf32.as<float>(), op->mf32.as<cx<float>>(), op->rsa32.as<float>(),
                           op->ma32.as<cx<float>>());
        HZ_LAUNCHED();
        HZ_HIP(hipDeviceSynchronize());
    });
    if (st != HZ_OK) {
        delete op;
        return st;
    }
    *out = op;
    return HZ_OK;
}

hz_status hz_destroy(hz_operator *op) {
    delete op;
    return HZ_OK;
}

hz_status hz_apply(hz_operator *op, int32_t precision, int32_t mode, const void *x, void *y, int32_t nrhs,
                   int64_t ld, void *stream) {
    return guarded([&] {
        check_op(op);
        if (!x || !y) throw BadArgument("x and y must be device pointers");
        if (nrhs < 1 || ld < op->n) throw BadArgument("need nrhs >= 1 and ld >= N");
        if (mode < MODE_A || mode > MODE_RAH) throw BadArgument("unknown apply mode");
        Fuser f(static_cast<hipStream_t>(stream));
        if (precision == HZ_C128)
            apply_mode<double>(f, 0, op, mode, static_cast<const cx<double> *>(x), ld, static_cast<cx<double> *>(y),
                               ld, nrhs, nullptr, nullptr);
        else if (precision == HZ_C64)
            apply_mode<float>(f, 0, op, mode, static_cast<const cx<float> *>(x), ld, static_cast<cx<float> *>(y), ld,
                              nrhs, nullptr, nullptr);
        else
            throw BadArgument("unknown precision");
    });
}

hz_status hz_gmres(hz_operator *op, const hz_gmres_opts *opts, const void *b, void *x, int32_t nrhs, int64_t ld,
                   hz_gmres_stats *stats, void *stream) {
    const hz_status st = guarded([&] { gmres_run(op, opts, b, x, nrhs, ld, stats, static_cast<hipStream_t>(stream)); });
    if (st != HZ_OK) return st;
    for (int r = 0; r < nrhs; ++r)
        if (!stats[r].converged) return HZ_NOT_CONVERGED;
    return HZ_OK;
}

hz_status hz_gmres_batch(int32_t nscenes, hz_operator *const *ops, const hz_gmres_opts *opts, const void *const *b,
                         void *const *x, const int32_t *nrhs, const int64_t *ld, hz_gmres_stats *const *stats,
                         int32_t workers, void *stream) {
    hipStream_t caller = static_cast<hipStream_t>(stream);
    std::vector<int> order;
    std::vector<hipStream_t> streams;
    hz_status st = guarded([&] {
        if (nscenes < 1 || !ops || !opts || !b || !x || !nrhs || !ld || !stats)
            throw BadArgument("need nscenes >= 1 and every per-scene array");
        for (int i = 0; i < nscenes; ++i) {
            check_op(ops[i]);
            for (int j = 0; j < i; ++j)
                if (ops[j] == ops[i]) throw BadArgument("an operator appears twice in one batch");
            order.push_back(i);
        }
        // Longest first, by nodes x right-hand sides.
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
            return (double)ops[a]->n * nrhs[a] > (double)ops[c]->n * nrhs[c];
        });
        const int nw = std::max(1, std::min({workers > 0 ? (int)workers : 4, (int)nscenes, 16}));
        hipEvent_t start;
        HZ_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
        HZ_HIP(hipEventRecord(start, caller));
        for (int w = 0; w < nw; ++w) {
            hipStream_t ws;
            HZ_HIP(hipStreamCreateWithFlags(&ws, hipStreamNonBlocking));
            streams.push_back(ws);
            HZ_HIP(hipStreamWaitEvent(ws, start, 0));
        }
        (void)hipEventDestroy(start);
    });
    if (st != HZ_OK) {
        for (hipStream_t ws : streams) (void)hipStreamDestroy(ws);
        return st;
    }
    int device = 0;
    (void)hipGetDevice(&device);
    std::vector<hz_status> status(nscenes, HZ_OK);
    std::vector<std::string> error(nscenes);
    std::atomic<int> next{0};
    auto worker = [&](hipStream_t ws) {
        (void)hipSetDevice(device);
        for (int idx = next++; idx < nscenes; idx = next++) {
            const int sc = order[idx];
            status[sc] = guarded([&] {
                gmres_run(ops[sc], opts, b[sc], x[sc], nrhs[sc], ld[sc], stats[sc], ws);
                // Bound the batch's memory to the scenes in flight: the solve
                // has synchronised its stream, so its workspace can go.
                HZ_HIP(hipStreamSynchronize(ws));
                drop_graphs(ops[sc]);
                release_workspace(ops[sc]);
            });
            if (status[sc] != HZ_OK) error[sc] = g_error;
        }
    };
    std::vector<std::thread> pool;
    for (size_t w = 1; w < streams.size(); ++w) pool.emplace_back(worker, streams[w]);
    worker(streams[0]);
    for (auto &t : pool) t.join();
    st = guarded([&] {
        for (hipStream_t ws : streams) {
            hipEvent_t done;
            HZ_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
            HZ_HIP(hipEventRecord(done, ws));
            HZ_HIP(hipStreamWaitEvent(caller, done, 0));
            (void)hipEventDestroy(done);
        }
    });
    for (hipStream_t ws : streams) {
        (void)hipStreamSynchronize(ws);
        (void)hipStreamDestroy(ws);
    }
    if (st != HZ_OK) return st;
    for (int sc = 0; sc < nscenes; ++sc)
        if (status[sc] != HZ_OK) {
            g_error = "scene " + std::to_string(sc) + ": " + error[sc];
            return status[sc];
        }
    for (int sc = 0; sc < nscenes; ++sc)
        for (int r = 0; r < nrhs[sc]; ++r)
            if (!stats[sc][r].converged) return HZ_NOT_CONVERGED;
    return HZ_OK;
}

hz_status hz_gmres_fused(int32_t nscenes, hz_operator *const *ops, const hz_gmres_opts *opts, const void *const *b,
                         void *const *x, const int32_t *nrhs, const int64_t *ld, hz_gmres_stats *const *stats,
                         void *stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const hz_status st = guarded([&] {
        if (nscenes < 1 || !ops || !opts || !b || !x || !nrhs || !ld || !stats)
            throw BadArgument("need nscenes >= 1 and every per-scene array");
        const hz_gmres_opts &o = *opts;
        if (o.precision != HZ_C64) throw BadArgument("fused multi-scene solves run in the mixed form (HZ_C64)");
        int most = 0;
        for (int i = 0; i < nscenes; ++i) {
            check_op(ops[i]);
            for (int j = 0; j < i; ++j)
                if (ops[j] == ops[i]) throw BadArgument("an operator appears twice in one batch");
            if (!b[i] || !x[i] || !stats[i]) throw BadArgument("b, x and stats are required for every scene");
            if (nrhs[i] < 1 || ld[i] < ops[i]->n) throw BadArgument("need nrhs >= 1 and ld >= N for every scene");
            check_opts(ops[i], o);
            most = std::max<int>(most, nrhs[i]);
        }
        // Every scene runs with the tile shapes of the largest scene that has
        // the level (results do not depend on the tile shape), so that the
        // scenes' launches at a position are one kernel and fuse; each
        // scene's own shapes are restored on return.
        struct Saved {
            int slab;
            std::vector<std::array<int, 3>> lv;
        };
        std::vector<Saved> saved(nscenes);
        auto lead_of = [&](int level) {
            int lead = -1;
            for (int i = 0; i < nscenes; ++i) {
                const int nl = ops[i]->mg ? (int)ops[i]->mg->levels.size() : 0;
                if ((level < 0 || level < nl) && (lead < 0 || ops[i]->n > ops[lead]->n)) lead = i;
            }
            return lead;
        };
        for (int i = 0; i < nscenes; ++i) {
            saved[i].slab = ops[i]->slab_shape;
            if (ops[i]->mg)
                for (auto &L : ops[i]->mg->levels) saved[i].lv.push_back({L->shape[0], L->shape[1], L->shape[2]});
        }
        struct Restore {
            hz_operator *const *ops;
            std::vector<Saved> &saved;
            ~Restore() {
                for (size_t i = 0; i < saved.size(); ++i) {
                    ops[i]->slab_shape = saved[i].slab;
                    for (size_t l = 0; l < saved[i].lv.size(); ++l)
                        for (int e = 0; e < 3; ++e) ops[i]->mg->levels[l]->shape[e] = saved[i].lv[l][e];
                }
            }
        } restore{ops, saved};
        const int lead = lead_of(-1);
        for (int i = 0; i < nscenes; ++i) {
            ops[i]->slab_shape = saved[lead].slab;
            for (size_t l = 0; l < saved[i].lv.size(); ++l) {
                const int src = lead_of((int)l);
                for (int e = 0; e < 3; ++e) ops[i]->mg->levels[l]->shape[e] = saved[src].lv[l][e];
            }
        }
        // Round q solves columns [q c, (q + 1) c) of every scene that has
        // them, all scenes in lockstep; the launch sequences (fused tables,
        // captured cycle bodies) are kept across the round's refinements.
        // The round width is opts->chunk, halved while the scenes' complex64
        // workspaces together would not fit in 90 % of the free device
        // memory.
        int chunk = chunk_width(o, most);
        {
            size_t free_b = 0, total_b = 0;
            HZ_HIP(hipMemGetInfo(&free_b, &total_b));
            auto need = [&](int c) {
                size_t bytes = 0;
                for (int i = 0; i < nscenes; ++i) bytes += chunk_bytes(ops[i], std::min(c, (int)nrhs[i]), o.restart);
                return bytes;
            };
            while (chunk > 1 && need(chunk) > free_b / 10 * 9) chunk = (chunk + 1) / 2;
        }
        LaunchCache cache;
        cache.limit = 64;
        for (int c0 = 0; c0 < most; c0 += chunk) {
            std::vector<MixedJob> jobs;
            for (int i = 0; i < nscenes; ++i) {
                if (c0 >= nrhs[i]) continue;
                MixedJob jb;
                jb.op = ops[i];
                jb.b = static_cast<const cx<double> *>(b[i]) + (int64_t)c0 * ld[i];
                jb.x = static_cast<cx<double> *>(x[i]) + (int64_t)c0 * ld[i];
                jb.ldb = jb.ldx = ld[i];
                jb.k = std::min(chunk, nrhs[i] - c0);
                jb.stats = stats[i] + c0;
                jobs.push_back(std::move(jb));
            }
            solve_mixed(jobs, o, s, &cache);
        }
        HZ_HIP(hipStreamSynchronize(s));
        cache.clear();
        for (int i = 0; i < nscenes; ++i) {
            drop_graphs(ops[i]);
            release_workspace(ops[i]);
        }
    });
    if (st != HZ_OK) return st;
    for (int sc = 0; sc < nscenes; ++sc)
        for (int r = 0; r < nrhs[sc]; ++r)
            if (!stats[sc][r].converged) return HZ_NOT_CONVERGED;
    return HZ_OK;
}

hz_status hz_launch_counts(int64_t *counts, int32_t reset) {
    if (!counts) {
        g_error = "counts is null";
        return HZ_BAD_ARGUMENT;
    }
    for (int q = 0; q < 3; ++q) counts[q] = reset ? g_launches[q].exchange(0) : g_launches[q].load();
    return HZ_OK;
}

hz_status hz_set_multigrid(hz_operator *op, int32_t nlevels, const hz_level_desc *levels, double shift, int32_t pre,
                           int32_t post, int32_t post_coarse, int32_t coarse_sweeps, double weight, int32_t cycle) {
    return guarded([&] {
        check_op(op);
        if (nlevels < 1 || !levels) throw BadArgument("need at least one level");
        if (pre < 1 || post < 0 || post_coarse < 0 || coarse_sweeps < 1 || !(weight > 0) || !std::isfinite(shift))
            throw BadArgument("bad multigrid smoothing parameters");
        if (cycle < 1 || cycle > 2) throw BadArgument("cycle must be 1 (V) or 2 (W)");
        auto mg = std::make_unique<Multigrid>();
        mg->cycle = cycle;
        mg->pre = pre;
        mg->post = post;
        mg->post_coarse = post_coarse;
        mg->coarse_sweeps = coarse_sweeps;
        mg->weight = (float)weight;
        mg->shift = shift;
        const double h2 = op->dx * op->dx;
        // Levels without host fields are derived on the device: 1/rho and
        // kappa of the level above (the operator's own at level 0) by
        // normalised full weighting, kept in FP64 until the level's
        // single-precision copies are written.
        const bool derive = !levels[0].rho && !levels[0].kappa;
        DevBuf cur_inv, cur_kap, nxt_inv, nxt_kap, t_inv, t_kap, t_w[2], t_s[2];
        for (int l = 0; l < nlevels; ++l) {
            const hz_level_desc &d = levels[l];
            const int dims[3] = {d.nx, d.ny, d.nz};
            if (derive != (!d.rho && !d.kappa))
                throw BadArgument("either every level gives rho and kappa or none does");
            if (l == 0 && (d.nx != op->nx || d.ny != op->ny || d.nz != op->nz))
                throw BadArgument("level 0 must be the operator's grid");
            if (l > 0) {
                const int prev[3] = {levels[l - 1].nx, levels[l - 1].ny, levels[l - 1].nz};
                for (int a = 0; a < 3; ++a) {
                    const int expect = (prev[a] - 1) / 2 + 1 + ((prev[a] - 1) % 2);
                    if (dims[a] != expect) throw BadArgument("level sizes must follow even-node coarsening");
                }
            }
            if (d.nx < 2 || d.ny < 2 || d.nz < 2) throw BadArgument("bad level grid");
            auto L = std::make_unique<MgLevel>();
            L->nx = d.nx;
            L->ny = d.ny;
            L->nz = d.nz;
            L->n = (int64_t)d.nx * d.ny * d.nz;
            if (derive) {
                const unsigned nb = (unsigned)((L->n + 255) / 256);
                if (l == 0) {
                    cur_inv.ensure(L->n * sizeof(double));
                    cur_kap.ensure(L->n * sizeof(double));
                    hipLaunchKernelGGL(k_mg_fields0, dim3(nb), dim3(256), 0, 0, op->rho64.as<double>(),
                                       op->kap64.as<double>(), L->n, cur_inv.as<double>(), cur_kap.as<double>());
                    HZ_LAUNCHED();
                } else {
                    // x, then y, then z: (px, py, pz) -> (nx, py, pz) -> (nx, ny, pz) -> (nx, ny, nz).
                    const int px = levels[l - 1].nx, py = levels[l - 1].ny, pz = levels[l - 1].nz;
                    const int64_t n1 = (int64_t)d.nx * py * pz, n2 = (int64_t)d.nx * d.ny * pz;
                    t_inv.ensure(n1 * sizeof(double));
                    t_kap.ensure(n1 * sizeof(double));
                    nxt_inv.ensure(n2 * sizeof(double));
                    nxt_kap.ensure(n2 * sizeof(double));
                    hipLaunchKernelGGL(k_mg_weight_axis, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, 0,
                                       cur_inv.as<double>(), cur_kap.as<double>(), px, py, pz, 0, d.nx,
                                       t_inv.as<double>(), t_kap.as<double>());
                    HZ_LAUNCHED();
                    hipLaunchKernelGGL(k_mg_weight_axis, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, 0,
                                       t_inv.as<double>(), t_kap.as<double>(), d.nx, py, pz, 1, d.ny,
                                       nxt_inv.as<double>(), nxt_kap.as<double>());
                    HZ_LAUNCHED();
                    hipLaunchKernelGGL(k_mg_weight_axis, dim3(nb), dim3(256), 0, 0, nxt_inv.as<double>(),
                                       nxt_kap.as<double>(), d.nx, d.ny, pz, 2, d.nz, cur_inv.as<double>(),
                                       cur_kap.as<double>());
                    HZ_LAUNCHED();
                }
                L->rho.ensure(L->n * sizeof(float));
                L->kap.ensure(L->n * sizeof(float));
                hipLaunchKernelGGL(k_mg_level_fields, dim3(nb), dim3(256), 0, 0, cur_inv.as<double>(),
                                   cur_kap.as<double>(), L->n, L->rho.as<float>(), L->kap.as<float>());
                HZ_LAUNCHED();
            } else {
                std::vector<float> f32((size_t)L->n);
                for (int64_t p = 0; p < L->n; ++p) f32[p] = (float)d.rho[p];
                upload(L->rho, f32.data(), L->n);
                for (int64_t p = 0; p < L->n; ++p) f32[p] = (float)d.kappa[p];
                upload(L->kap, f32.data(), L->n);
            }
            for (int a = 0; a < 3; ++a) {
                const int na = dims[a];
                if (!d.weight[a] || !d.stretch[a] || !d.edge[a]) throw BadArgument("level tables are required");
                std::vector<cx<float>> av(na), qv(std::max(na - 1, 1));
                for (int i = 0; i < na; ++i)
                    av[i] = cast<float>(d.weight[a][i] * mk<double>(d.stretch[a][2 * i], d.stretch[a][2 * i + 1]));
                for (int i = 0; i + 1 < na; ++i) {
                    const cx<double> s0 = mk<double>(d.stretch[a][2 * i], d.stretch[a][2 * i + 1]);
                    const cx<double> s1 = mk<double>(d.stretch[a][2 * i + 2], d.stretch[a][2 * i + 3]);
                    qv[i] = cast<float>(4.0 * crecip((d.edge[a][i] * h2) * (s0 + s1)));
                }
                upload(L->a[a], av.data(), na);
                upload(L->q[a], qv.data(), qv.size());
            }
            for (int f = 0; f < 6; ++f) {
                const int ax = f / 2;
                const int b0 = ax == 0 ? 1 : 0, c0 = ax == 2 ? 1 : 2;
                const int64_t cnt = (int64_t)dims[b0] * dims[c0];
                if (!d.face_term[f] && derive && (op->open_mask >> f & 1)) {
                    // Open face of a derived level: its term from the derived fields.
                    const int ab[2] = {b0, c0};
                    for (int u = 0; u < 2; ++u) {
                        upload(t_w[u], d.weight[ab[u]], (size_t)dims[ab[u]]);
                        upload(t_s[u], reinterpret_cast<const cx<double> *>(d.stretch[ab[u]]), (size_t)dims[ab[u]]);
                    }
                    L->face[f].ensure(cnt * sizeof(cx<float>));
                    hipLaunchKernelGGL(k_mg_open_face, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, 0,
                                       cur_inv.as<double>(), cur_kap.as<double>(), d.nx, d.ny, d.nz, f,
                                       t_w[0].as<double>(), t_s[0].as<cx<double>>(), t_w[1].as<double>(),
                                       t_s[1].as<cx<double>>(), op->omega, op->dx, L->face[f].as<cx<float>>());
                    HZ_LAUNCHED();
                    // The next face reuses t_w / t_s: wait for this launch.
                    HZ_HIP(hipDeviceSynchronize());
                    continue;
                }
                if (!d.face_term[f]) continue;
                std::vector<cx<float>> t32(cnt);
                for (int64_t e = 0; e < cnt; ++e)
                    t32[e] = mk<float>((float)d.face_term[f][2 * e], (float)d.face_term[f][2 * e + 1]);
                upload(L->face[f], t32.data(), cnt);
            }
            OpView<float> &v = L->v;
            v.nx = L->nx;
            v.ny = L->ny;
            v.nz = L->nz;
            v.n = L->n;
            v.omega2 = (float)(op->omega * op->omega);
            v.omega2_im = (float)(shift * op->omega * op->omega);
            v.rho = L->rho.as<float>();
            v.kappa = L->kap.as<float>();
            for (int a = 0; a < 3; ++a) {
                v.a[a] = L->a[a].as<cx<float>>();
                v.q[a] = L->q[a].as<cx<float>>();
            }
            for (int f = 0; f < 6; ++f) v.face[f] = L->face[f].p ? L->face[f].as<cx<float>>() : nullptr;
            v.soft = op->soft;
            v.rs_f = v.rs_a = nullptr;
            v.m_f = v.m_a = nullptr;
            L->tm = make_tilemap(L->nx, L->ny, L->nz, 4);
            L->dinv.ensure(L->n * sizeof(cx<float>));
            hipLaunchKernelGGL((k_level_dinv<float>), dim3((L->n + 255) / 256), dim3(256), 0, 0, L->v,
                               L->dinv.as<cx<float>>());
            HZ_LAUNCHED();
            mg->levels.push_back(std::move(L));
        }
        HZ_HIP(hipDeviceSynchronize());
        op->mg = std::move(mg);
    });
}

hz_status hz_set_sources(hz_operator *op, int32_t num_columns, int64_t nnz, const int64_t *node,
                         const int32_t *column, const double *value) {
    return guarded([&] {
        check_op(op);
        if (num_columns < 1 || nnz < 1 || !node || !column || !value) throw BadArgument("empty source list");
        for (int64_t e = 0; e < nnz; ++e)
            if (node[e] < 0 || node[e] >= op->n || column[e] < 0 || column[e] >= num_columns)
                throw BadArgument("source entry out of range");
        std::vector<int64_t> order(nnz);
        for (int64_t e = 0; e < nnz; ++e) order[e] = e;
        // Grouped by node (scatter).
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            return node[a] != node[b] ? node[a] < node[b] : column[a] < column[b];
        });
        std::vector<int64_t> rnode, rptr{0};
        std::vector<int> rcol(nnz);
        std::vector<double> rval(nnz);
        for (int64_t e = 0; e < nnz; ++e) {
            const int64_t src = order[e];
            if (e == 0 || node[src] != node[order[e - 1]]) {
                if (e) rptr.push_back(e);
                rnode.push_back(node[src]);
            }
            rcol[e] = column[src];
            rval[e] = value[src];
        }
        rptr.push_back(nnz);
        // Grouped by column (gather).
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            return column[a] != column[b] ? column[a] < column[b] : node[a] < node[b];
        });
        std::vector<int64_t> cnode(nnz), cptr(num_columns + 1, 0);
        std::vector<double> cval(nnz);
        for (int64_t e = 0; e < nnz; ++e) {
            cnode[e] = node[order[e]];
            cval[e] = value[order[e]];
            cptr[column[order[e]] + 1] += 1;
        }
        for (int c = 0; c < num_columns; ++c) cptr[c + 1] += cptr[c];
        upload(op->src_row_node, rnode.data(), rnode.size());
        upload(op->src_row_ptr, rptr.data(), rptr.size());
        upload(op->src_row_col, rcol.data(), rcol.size());
        upload(op->src_row_val, rval.data(), rval.size());
        upload(op->src_col_node, cnode.data(), cnode.size());
        upload(op->src_col_ptr, cptr.data(), cptr.size());
        upload(op->src_col_val, cval.data(), cval.size());
        op->ncols = num_columns;
        op->nrows_src = (int64_t)rnode.size();
    });
}

hz_status hz_build_rhs(hz_operator *op, const void *w, int64_t ld_w, void *b, int32_t nrhs, int64_t ld,
                       void *stream) {
    return guarded([&] {
        check_op(op);
        if (op->ncols == 0) throw BadArgument("no source columns set");
        if (!w || !b || nrhs < 1 || ld < op->n || ld_w < op->ncols) throw BadArgument("bad rhs arguments");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL((k_zero<double>), dim3(vec_blocks(op->n), nrhs), dim3(256), 0, s,
                           static_cast<cx<double> *>(b), ld, op->n);
        HZ_LAUNCHED();
        const int64_t blocks = blocks_for(op->nrows_src, 256, 4096);
        hipLaunchKernelGGL(k_scatter_sources, dim3(blocks, nrhs), dim3(256), 0, s, op->src_row_node.as<int64_t>(),
                           op->src_row_ptr.as<int64_t>(), op->src_row_col.as<int>(), op->src_row_val.as<double>(),
                           op->nrows_src, static_cast<const cx<double> *>(w), ld_w, static_cast<cx<double> *>(b), ld);
        HZ_LAUNCHED();
    });
}

hz_status hz_gather_sources(hz_operator *op, const void *lambda, int64_t ld, void *out, int64_t ld_out,
                            int32_t nrhs, void *stream) {
    return guarded([&] {
        check_op(op);
        if (op->ncols == 0) throw BadArgument("no source columns set");
        if (!lambda || !out || nrhs < 1 || ld < op->n || ld_out < op->ncols) throw BadArgument("bad gather arguments");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_gather_sources, dim3(op->ncols, nrhs), dim3(256), 0, s, op->src_col_node.as<int64_t>(),
                           op->src_col_val.as<double>(), op->src_col_ptr.as<int64_t>(),
                           static_cast<const cx<double> *>(lambda), ld, static_cast<cx<double> *>(out), ld_out);
        HZ_LAUNCHED();
    });
}

hz_status hz_sample(hz_operator *op, const int64_t *rows, const double *weights, int32_t npoints, const void *u,
                    int64_t ld, void *out, int64_t ld_out, int32_t nrhs, void *stream) {
    return guarded([&] {
        check_op(op);
        if (!rows || !weights || !u || !out || npoints < 1 || nrhs < 1 || ld < op->n || ld_out < npoints)
            throw BadArgument("bad sample arguments");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_sample, dim3((npoints + 255) / 256, nrhs), dim3(256), 0, s, rows, weights, npoints,
                           static_cast<const cx<double> *>(u), ld, static_cast<cx<double> *>(out), ld_out);
        HZ_LAUNCHED();
    });
}

hz_status hz_autotune(hz_operator *op, int32_t nrhs, int32_t max_chunk, int32_t restart, int32_t preconditioner,
                      int64_t memory_budget, int32_t *chunk_out, int32_t *shapes_out, int32_t shapes_len,
                      void *stream) {
    return guarded([&] {
        check_op(op);
        if (nrhs < 1 || max_chunk < 1) throw BadArgument("need nrhs >= 1 and max_chunk >= 1");
        if (restart < 1 || restart > 63) throw BadArgument("restart must be in [1, 63]");
        const bool use_mg = preconditioner == PRE_MULTIGRID;
        if (use_mg && !op->mg) throw BadArgument("multigrid preconditioner requested but no hierarchy installed");
        if (!use_mg && preconditioner != PRE_JACOBI) throw BadArgument("unknown preconditioner");
        hipStream_t s = static_cast<hipStream_t>(stream);
        Fuser fs(s);
        const int64_t n = op->n;
        size_t free_b = 0, total_b = 0;
        HZ_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t budget = memory_budget > 0 ? (size_t)memory_budget : free_b / 2;
        // Candidate chunk widths that fit the budget.
        std::vector<int> cand;
        for (int c : {4, 8, 16, 32, 64})
            if (c <= max_chunk && c <= std::max(nrhs, 4) && chunk_bytes(op, c, restart) <= budget) cand.push_back(c);
        if (cand.empty()) cand.push_back(std::max(1, std::min({max_chunk, nrhs, 4})));
        // Columns for the shape timings: up to 32.  At 16 the paired shapes
        // with 2 and 4 z-planes per lane tie; from 32 on (the widths the
        // solve runs) the 4-plane one, which re-fetches fewer halo planes,
        // is 5-6 % faster on the 128^3 apply.
        const int kt = std::min(32, cand.back());
        const int kmax = cand.back();
        reserve_workspace(op, kmax, restart, sizeof(cx<float>));
        if (use_mg) mg_reserve(*op->mg, kmax);
        // Finite, non-zero data: the basis and iterate start as ones; unit scales.
        {
            std::vector<int> ones_i(kmax, 1);
            std::vector<double> ones_d((size_t)(restart + 1) * kmax, 1.0);
            std::vector<cx<double>> coef((size_t)kmax * (restart + 1), mk<double>(1e-3, 0));
            HZ_HIP(hipMemcpyAsync(op->live.p, ones_i.data(), kmax * sizeof(int), hipMemcpyHostToDevice, s));
            HZ_HIP(hipMemcpyAsync(op->scales.p, ones_d.data(), ones_d.size() * sizeof(double), hipMemcpyHostToDevice,
                                  s));
            HZ_HIP(hipMemcpyAsync(op->coef.p, coef.data(), coef.size() * sizeof(cx<double>), hipMemcpyHostToDevice,
                                  s));
            hipLaunchKernelGGL((k_fill<float>), dim3(vec_blocks(n), (unsigned)((restart + 1) * kmax)), dim3(256), 0, s,
                               op->V.as<cx<float>>(), n, n, mk<float>(1, 0));
            HZ_LAUNCHED();
            if (use_mg)
                for (auto &L : op->mg->levels)
                    for (DevBuf *b : {&L->bb, &L->xx, &L->tt}) {
                        hipLaunchKernelGGL((k_fill<float>), dim3(vec_blocks(L->n), kmax), dim3(256), 0, s,
                                           b->as<cx<float>>(), L->n, L->n, mk<float>(1, 0));
                        HZ_LAUNCHED();
                    }
        }
        const int *live = op->live.as<int>();
        cx<float> *x = op->V.as<cx<float>>();
        cx<float> *y = x + (int64_t)kt * n;
        // Fine-grid apply.
        const int mode = use_mg ? MODE_RA : MODE_RAM;
        warm_up(
            [&] {
                launch_slab<MODE_A, SLAB_APPLY>(fs, 0, op->slab_shape, op->v32, x, n, y, n, kt, live,
                                                op->scales.as<double>(), nullptr, nullptr, 0.0f, nullptr);
            },
            0.3, s);
        op->slab_shape = pick_shape(
            [&](int sh) {
                if (mode == MODE_RA)
                    launch_slab<MODE_RA, SLAB_APPLY>(fs, 0, sh, op->v32, x, n, y, n, kt, live, op->scales.as<double>(),
                                                     nullptr, nullptr, 0.0f, nullptr);
                else
                    launch_slab<MODE_RAM, SLAB_APPLY>(fs, 0, sh, op->v32, x, n, y, n, kt, live, op->scales.as<double>(),
                                                      nullptr, nullptr, 0.0f, nullptr);
            },
            s);
        // Multigrid level kernels (forward operator; the adjoint uses the same shapes).
        if (use_mg) {
            Multigrid &mg = *op->mg;
            for (size_t l = 0; l + 1 < mg.levels.size() || (l < mg.levels.size() && mg.levels[l]->n > COARSE_MAX_NODES);
                 ++l) {
                MgLevel &L = *mg.levels[l];
                const cx<float> *lb = L.bb.as<cx<float>>();
                cx<float> *lx = L.xx.as<cx<float>>(), *lt = L.tt.as<cx<float>>();
                const cx<float> *dv = L.dinv.as<cx<float>>();
                const float *bd = l == 0 ? op->v32.rs_f : nullptr;
                L.shape[0] = pick_shape(
                    [&](int sh) {
                        launch_slab<MODE_A, SLAB_SMOOTH>(fs, 0, sh, L.v, lx, L.n, lt, L.n, kt, live, nullptr, lb, dv,
                                                         mg.weight, bd);
                    },
                    s);
                L.shape[1] = pick_shape(
                    [&](int sh) {
                        launch_slab<MODE_A, SLAB_RESID>(fs, 0, sh, L.v, lx, L.n, lt, L.n, kt, live, nullptr, lb,
                                                        nullptr, 0.0f, bd);
                    },
                    s);
                L.shape[2] = pick_shape(
                    [&](int sh) {
                        launch_slab<MODE_A, SLAB_RESID0>(fs, 0, sh, L.v, lb, L.n, lt, L.n, kt, live, nullptr,
                                                         nullptr, dv, mg.weight, bd);
                    },
                    s, 1);
            }
        }
        // Chunk width: least time per column of one Arnoldi step at j = restart / 2.
        // Widths within 2 % of the least go to the widest of them: the step
        // timing does not see the per-cycle costs (cycle tops, refinement
        // residuals, host round trips) that a wider chunk spreads over more
        // columns.  The memory budget already bounds the candidates.
        std::vector<double> per_col(cand.size());
        double least = 0;
        for (size_t ci = 0; ci < cand.size(); ++ci) {
            const int c = cand[ci];
            const float ms = time_launches([&] { arnoldi_step_for_timing(op, c, restart / 2, restart, use_mg, s); },
                                           3, s);
            per_col[ci] = ms / c;
            if (verbose()) std::fprintf(stderr, "[hz] autotune chunk %d: %.3f ms/step, %.1f us per column\n", c, ms,
                                        per_col[ci] * 1e3);
            if (ci == 0 || per_col[ci] < least) least = per_col[ci];
        }
        int best = cand.front();
        for (size_t ci = 0; ci < cand.size(); ++ci)
            if (per_col[ci] <= 1.02 * least) best = cand[ci];
        HZ_HIP(hipStreamSynchronize(s));
        release_workspace(op);
        if (chunk_out) *chunk_out = best;
        if (shapes_out && shapes_len > 0) {
            shapes_out[0] = op->slab_shape;
            int at = 1;
            if (op->mg)
                for (const auto &L : op->mg->levels)
                    for (int e = 0; e < 3 && at < shapes_len; ++e) shapes_out[at++] = L->shape[e];
        }
    });
}

hz_status hz_basis_field(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w, void *u,
                         void *stream) {
    return guarded([&] {
        check_basis_args(G, ldg, nrows, m);
        if (!w || !u) throw BadArgument("w and u must be device pointers");
        const int nb = (int)std::min<int64_t>(4096, (nrows + 3) / 4);
        hipLaunchKernelGGL(k_basis_field, dim3(nb), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const cx<double> *>(G), ldg, nrows, (int)m, static_cast<const cx<double> *>(w),
                           static_cast<cx<double> *>(u));
        HZ_LAUNCHED();
    });
}

}  // extern "C"

namespace {

// Workgroups of a fused loss launch (4 waves each): at most 1024, and their
// partials plus the folded slices within BASIS_LOSS_PARTIAL_BYTES.
int64_t basis_loss_blocks(int64_t nrows, int m, int nacc) {
    const int64_t per = (int64_t)nacc * m * (int64_t)sizeof(cx<double>);
    const int64_t cap = (BASIS_LOSS_PARTIAL_BYTES - LOSS_FOLD_SLICES * per) / per;
    return std::max<int64_t>(1, std::min<int64_t>({(int64_t)1024, cap, (nrows + 3) / 4}));
}

// Slices the workgroup partials fold into (0: combined directly); the
// combine takes at most LOSS_FOLD_SLICES partials and a slice at most 64
// (1024 workgroups / 64 slices = 16).
int64_t basis_loss_slices(int64_t nblk) { return nblk > LOSS_FOLD_SLICES ? LOSS_FOLD_SLICES : 0; }

BasisLossArgs basis_loss_args(const hz_basis_loss_desc *d, int64_t nrows) {
    if (!d) throw BadArgument("the loss descriptor is required");
    BasisLossArgs a{};
    a.target = d->target;
    a.weight = d->weight;
    a.mask = d->mask;
    a.eps = d->eps;
    a.scale = d->scale;
    a.temperature = d->temperature;
    a.count_a = (double)d->count_a;
    a.count_b = (double)d->count_b;
    switch (d->kind) {
    case HZ_LOSS_FIELD_MATCH:
        if (!a.target || !a.weight) throw BadArgument("field_match needs target and weight");
        break;
    case HZ_LOSS_ENERGY:
        break;
    case HZ_LOSS_SOFTMAX:
        if (!a.mask) throw BadArgument("the smooth-max loss needs a row mask");
        if (!(a.temperature > 0)) throw BadArgument("the smooth-max temperature must be positive");
        if (d->count_a < 1 || d->count_a > nrows) throw BadArgument("count_a must be the number of masked rows");
        break;
    case HZ_LOSS_RATIO:
        if (!a.mask) throw BadArgument("the ratio loss needs a row class mask");
        if (d->count_a < 1 || d->count_b < 1 || d->count_a + d->count_b > nrows)
            throw BadArgument("count_a and count_b must be the sizes of row classes 1 and 2");
        break;
    default:
        throw BadArgument("unknown loss kind");
    }
    return a;
}

template <int KIND, int PER>
void launch_basis_loss_per(const cx<double> *G, int64_t ldg, int64_t nrows, int m, const cx<double> *w,
                           const BasisLossArgs &a, int64_t nblk, cx<double> *part, double *scal, cx<double> *u,
                           hipStream_t s) {
    const int64_t rows = (nrows + 4 * nblk - 1) / (4 * nblk);
    hipLaunchKernelGGL((k_basis_loss<KIND, PER>), dim3((unsigned)nblk), dim3(256), 0, s, G, ldg, nrows, m, w, a, rows,
                       part, scal, u);
    HZ_LAUNCHED();
}

// Scratch layout: workgroup partials, folded partials, workgroup scalars,
// folded scalars.
template <int KIND>
void launch_basis_loss_kind(const cx<double> *G, int64_t ldg, int64_t nrows, int m, const cx<double> *w,
                            const BasisLossArgs &a, void *scratch_base, double *loss, cx<double> *g, cx<double> *u,
                            hipStream_t s) {
    constexpr int nacc = KIND == LOSS_RATIO ? 2 : 1;
    const int64_t nblk = basis_loss_blocks(nrows, m, nacc);
    const int64_t nsl = basis_loss_slices(nblk);
    cx<double> *part = static_cast<cx<double> *>(scratch_base);
    cx<double> *fold = part + (int64_t)nacc * nblk * m;
    double *scal = reinterpret_cast<double *>(fold + (int64_t)nacc * nsl * m);
    double *fscal = scal + nblk * LOSS_SCALARS;
    if (m <= 128) launch_basis_loss_per<KIND, 2>(G, ldg, nrows, m, w, a, nblk, part, scal, u, s);
    else if (m <= 256) launch_basis_loss_per<KIND, 4>(G, ldg, nrows, m, w, a, nblk, part, scal, u, s);
    else launch_basis_loss_per<KIND, 8>(G, ldg, nrows, m, w, a, nblk, part, scal, u, s);
    int64_t nparts = nblk;
    if (nsl) {
        hipLaunchKernelGGL(k_basis_loss_fold<KIND>, dim3((unsigned)((m + 63) / 64), (unsigned)nsl), dim3(64), 0, s,
                           (const cx<double> *)part, (const double *)scal, nblk, m, a.temperature, fold, fscal, nsl);
        HZ_LAUNCHED();
        part = fold;
        scal = fscal;
        nparts = nsl;
    }
    hipLaunchKernelGGL(k_basis_loss_combine<KIND>, dim3((m + 63) / 64), dim3(256), 0, s,
                       (const cx<double> *)part, (const double *)scal, nparts, m, a, g, loss);
    HZ_LAUNCHED();
}

// One fused loss evaluation with its partials at `scratch_base` (which holds
// at least basis_loss_bytes(...) bytes).
void launch_basis_loss(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                       const BasisLossArgs &a, int kind, void *scratch_base, double *loss, void *g, void *u,
                       hipStream_t s) {
    auto Gc = static_cast<const cx<double> *>(G);
    auto wc = static_cast<const cx<double> *>(w);
    auto gc = static_cast<cx<double> *>(g);
    auto uc = static_cast<cx<double> *>(u);
    switch (kind) {
    case HZ_LOSS_FIELD_MATCH:
        launch_basis_loss_kind<LOSS_FIELD_MATCH>(Gc, ldg, nrows, m, wc, a, scratch_base, loss, gc, uc, s);
        break;
    case HZ_LOSS_ENERGY:
        launch_basis_loss_kind<LOSS_ENERGY>(Gc, ldg, nrows, m, wc, a, scratch_base, loss, gc, uc, s);
        break;
    case HZ_LOSS_SOFTMAX:
        launch_basis_loss_kind<LOSS_SOFTMAX>(Gc, ldg, nrows, m, wc, a, scratch_base, loss, gc, uc, s);
        break;
    default:
        launch_basis_loss_kind<LOSS_RATIO>(Gc, ldg, nrows, m, wc, a, scratch_base, loss, gc, uc, s);
        break;
    }
}

size_t basis_loss_bytes(int64_t nrows, int32_t m, int kind) {
    const int nacc = kind == HZ_LOSS_RATIO ? 2 : 1;
    const int64_t nblk = basis_loss_blocks(nrows, m, nacc);
    const int64_t nsl = basis_loss_slices(nblk);
    return (size_t)nacc * (nblk + nsl) * m * sizeof(cx<double>) + (size_t)(nblk + nsl) * LOSS_SCALARS * sizeof(double);
}

// G^H c for a given cotangent (no loss) through the fused kernels.
void launch_basis_adjoint(const cx<double> *G, int64_t ldg, int64_t nrows, int m, const BasisLossArgs &a,
                          void *scratch_base, cx<double> *g, hipStream_t s) {
    launch_basis_loss_kind<LOSS_ADJOINT>(G, ldg, nrows, m, nullptr, a, scratch_base, nullptr, g, nullptr, s);
}

void check_basis_loss_args(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w, double *loss,
                           void *g) {
    check_basis_args(G, ldg, nrows, m);
    if (m > BASIS_FUSED_MAX_M) throw BadArgument("the fused loss kernels take at most 512 emitters");
    if (!w || !loss || !g) throw BadArgument("w, loss and g must be device pointers");
}

}  // namespace

extern "C" {

hz_status hz_basis_adjoint(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *c, void *g,
                           void *stream) {
    return guarded([&] {
        check_basis_args(G, ldg, nrows, m);
        if (!c || !g) throw BadArgument("c and g must be device pointers");
        BasisLossArgs a{};
        a.cot = static_cast<const cx<double> *>(c);
        // Columns in slices of at most BASIS_FUSED_MAX_M (each reads its own
        // columns of G once); the slices share the workspace in stream order.
        const int32_t width = std::min<int32_t>(m, BASIS_FUSED_MAX_M);
        DevBuf &ws = basis_workspace(basis_loss_bytes(nrows, width, HZ_LOSS_ENERGY));
        for (int32_t c0 = 0; c0 < m; c0 += BASIS_FUSED_MAX_M)
            launch_basis_adjoint(static_cast<const cx<double> *>(G) + c0, ldg, nrows,
                                 std::min<int32_t>(BASIS_FUSED_MAX_M, m - c0), a, ws.p,
                                 static_cast<cx<double> *>(g) + c0, static_cast<hipStream_t>(stream));
    });
}

hz_status hz_basis_field_match(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                               const double *target, const double *weight, double eps, double *loss, void *g, void *u,
                               void *stream) {
    return guarded([&] {
        check_basis_loss_args(G, ldg, nrows, m, w, loss, g);
        hz_basis_loss_desc d{};
        d.kind = HZ_LOSS_FIELD_MATCH;
        d.target = target;
        d.weight = weight;
        d.eps = eps;
        const BasisLossArgs a = basis_loss_args(&d, nrows);
        DevBuf &ws = basis_workspace(basis_loss_bytes(nrows, m, d.kind));
        launch_basis_loss(G, ldg, nrows, m, w, a, d.kind, ws.p, loss, g, u, static_cast<hipStream_t>(stream));
    });
}

hz_status hz_basis_loss(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                        const hz_basis_loss_desc *desc, double *loss, void *g, void *u, void *stream) {
    return guarded([&] {
        check_basis_loss_args(G, ldg, nrows, m, w, loss, g);
        const BasisLossArgs a = basis_loss_args(desc, nrows);
        DevBuf &ws = basis_workspace(basis_loss_bytes(nrows, m, desc->kind));
        launch_basis_loss(G, ldg, nrows, m, w, a, desc->kind, ws.p, loss, g, u, static_cast<hipStream_t>(stream));
    });
}

hz_status hz_phase_prologue(const double *phases, int32_t m, const void *response, int32_t ntargets, void *w,
                            void *focal, void *stream) {
    return guarded([&] {
        if (!phases || !w) throw BadArgument("phases and w must be device pointers");
        if (m < 1 || m > BASIS_FUSED_MAX_M) throw BadArgument("need 1 <= m <= 512 emitters");
        if (ntargets < 0 || ntargets > PHASE_MAX_TARGETS) throw BadArgument("need 0 <= ntargets <= 256");
        if (ntargets > 0 && (!response || !focal)) throw BadArgument("targets need the response and focal buffers");
        hipLaunchKernelGGL(k_phase_prologue, dim3(1), dim3(256), 0, static_cast<hipStream_t>(stream), phases, (int)m,
                           static_cast<const cx<double> *>(response), (int)ntargets, static_cast<cx<double> *>(w),
                           static_cast<cx<double> *>(focal));
        HZ_LAUNCHED();
    });
}

hz_status hz_phase_epilogue(const void *w, int32_t m, const void *g, const void *response, int32_t ntargets,
                            const void *focal, double target_weight, double eps, const double *loss,
                            int32_t with_gradient, double *out, void *stream) {
    return guarded([&] {
        if (!w || !loss || !out) throw BadArgument("w, loss and out must be device pointers");
        if (m < 1 || m > BASIS_FUSED_MAX_M) throw BadArgument("need 1 <= m <= 512 emitters");
        if (ntargets < 0 || ntargets > PHASE_MAX_TARGETS) throw BadArgument("need 0 <= ntargets <= 256");
        if (ntargets > 0 && (!response || !focal)) throw BadArgument("targets need the response and focal buffers");
        if (with_gradient && !g) throw BadArgument("the gradient needs g = G^H c");
        if (!(eps > 0) || !std::isfinite(target_weight)) throw BadArgument("need eps > 0 and a finite target weight");
        hipLaunchKernelGGL(k_phase_epilogue, dim3(1), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const cx<double> *>(w), (int)m, static_cast<const cx<double> *>(g),
                           static_cast<const cx<double> *>(response), (int)ntargets,
                           static_cast<const cx<double> *>(focal), target_weight, eps, loss, (int)(with_gradient != 0),
                           out);
        HZ_LAUNCHED();
    });
}

}  // extern "C"

// Scratch of hz_basis_loss_static: the per-workgroup and folded partials of
// one fused loss launch (at most BASIS_LOSS_PARTIAL_BYTES of accumulators plus
// two doubles for each of at most 1024 + 64 partials; 4096 reserved), in
// device memory that exists from module load to process end, one copy per GPU.  Its host-side name is not a device
// address: host code uses only the pointer hipGetSymbolAddress returns.
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)BASIS_LOSS_PARTIAL_BYTES + 4096 * LOSS_SCALARS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address).
void *static_scratch() {
    static std::mutex mu;
    static std::map<int, void *> ptrs;
    int dev = 0;
    HZ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    void *ptr = nullptr;
    HZ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw HipFailure(hipErrorInvalidSymbol, "hipGetSymbolAddress returned a null scratch address");
    ptrs[dev] = ptr;
    return ptr;
}

}  // namespace

extern "C" {

hz_status hz_basis_loss_static(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                               const hz_basis_loss_desc *desc, double *loss, void *g, void *u, void *stream) {
    return guarded([&] {
        check_basis_loss_args(G, ldg, nrows, m, w, loss, g);
        const BasisLossArgs a = basis_loss_args(desc, nrows);
        const size_t need = basis_loss_bytes(nrows, m, desc->kind);
        if (need > STATIC_SCRATCH_BYTES)
            throw BadArgument("the partial sums need " + std::to_string(need) + " bytes, more than the " +
                              std::to_string(STATIC_SCRATCH_BYTES) + "-byte static scratch");
        // Every launch writes each partial and partial scalar it later reads, so
        // nothing in the scratch needs clearing between calls.
        launch_basis_loss(G, ldg, nrows, m, w, a, desc->kind, static_scratch(), loss, g, u,
                          static_cast<hipStream_t>(stream));
    });
}

hz_status hz_material_vjp(hz_operator *op, const void *u, const void *lambda, int32_t nrhs, int64_t ld, double *g_rho,
                          double *g_c, void *stream) {
    return guarded([&] {
        check_op(op);
        if (!u || !lambda || !g_rho || !g_c) throw BadArgument("u, lambda, g_rho and g_c must be device pointers");
        if (nrhs < 1 || ld < op->n) throw BadArgument("need nrhs >= 1 and ld >= N");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_material_vjp, dim3(blocks_for(op->n, 256, 65536)), dim3(256), 0, s, op->v64,
                           op->open_mask, static_cast<const cx<double> *>(u), static_cast<const cx<double> *>(lambda),
                           ld, (int)nrhs, g_rho, g_c);
        HZ_LAUNCHED();
    });
}

hz_status hz_release_workspace(hz_operator *op) {
    return guarded([&] {
        check_op(op);
        HZ_HIP(hipDeviceSynchronize());
        drop_graphs(op);
        release_workspace(op);
    });
}

hz_status hz_mg_transfer(hz_operator *op, int32_t level, int32_t direction, const void *in, void *out, int32_t nrhs,
                         void *stream) {
    return guarded([&] {
        check_op(op);
        if (!op->mg) throw BadArgument("no multigrid hierarchy set (hz_set_multigrid)");
        const int levels = (int)op->mg->levels.size();
        if (level < 0 || level + 1 >= levels) throw BadArgument("level must name a level with a coarser one below it");
        if (direction != 0 && direction != 1) throw BadArgument("direction must be 0 (restrict) or 1 (prolong)");
        if (!in || !out || nrhs < 1) throw BadArgument("in, out and nrhs >= 1 are required");
        const MgLevel &L = *op->mg->levels[level], &C = *op->mg->levels[level + 1];
        auto a16 = [](const void *ptr) { return ((uintptr_t)ptr & 15) == 0; };
        if (L.nx % 2 == 0 && !a16(direction == 0 ? in : out))
            throw BadArgument("even-width fine vectors must be 16-byte aligned");
        Fuser f(static_cast<hipStream_t>(stream));
        if (direction == 0)
            launch_restrict(f, 0, L, C, static_cast<const cx<float> *>(in), static_cast<cx<float> *>(out), nrhs,
                            nullptr);
        else
            launch_prolong(f, 0, L, C, static_cast<const cx<float> *>(in), static_cast<cx<float> *>(out), nrhs,
                           nullptr, nullptr, nullptr, 0.0f, nullptr, 0);
    });
}

hz_status hz_mg_level(hz_operator *op, int32_t level, float *rho, float *kappa, void *const *face,
                      int32_t *face_mask, void *stream) {
    return guarded([&] {
        check_op(op);
        if (!op->mg) throw BadArgument("no multigrid hierarchy is installed");
        if (level < 0 || level >= (int)op->mg->levels.size()) throw BadArgument("level out of range");
        if (!face_mask) throw BadArgument("face_mask is required");
        const MgLevel &L = *op->mg->levels[level];
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (rho) HZ_HIP(hipMemcpyAsync(rho, L.rho.p, L.n * sizeof(float), hipMemcpyDeviceToDevice, s));
        if (kappa) HZ_HIP(hipMemcpyAsync(kappa, L.kap.p, L.n * sizeof(float), hipMemcpyDeviceToDevice, s));
        const int dims[3] = {L.nx, L.ny, L.nz};
        int mask = 0;
        for (int f = 0; f < 6; ++f) {
            if (!L.face[f].p) continue;
            mask |= 1 << f;
            const int ax = f / 2;
            const int64_t cnt = (int64_t)dims[ax == 0 ? 1 : 0] * dims[ax == 2 ? 1 : 2];
            if (face && face[f])
                HZ_HIP(hipMemcpyAsync(face[f], L.face[f].p, cnt * sizeof(cx<float>), hipMemcpyDeviceToDevice, s));
        }
        *face_mask = mask;
    });
}

hz_status hz_set_tile_shapes(hz_operator *op, const int32_t *shapes, int32_t len) {
    return guarded([&] {
        check_op(op);
        const int levels = op->mg ? (int)op->mg->levels.size() : 0;
        if (!shapes || len != 1 + 3 * levels) throw BadArgument("need 1 + 3 * (multigrid levels) shape indices");
        for (int i = 0; i < len; ++i)
            if (shapes[i] < 0 || shapes[i] >= SLAB_SHAPES) throw BadArgument("shape index out of range");
        op->slab_shape = shapes[0];
        for (int l = 0; l < levels; ++l)
            for (int e = 0; e < 3; ++e) op->mg->levels[l]->shape[e] = shapes[1 + 3 * l + e];
    });
}

int64_t hz_device_bytes(const hz_operator *op) {
    if (!op) return 0;
    int64_t total = 0;
    const DevBuf *bufs[] = {&op->rho64, &op->kap64, &op->rho32, &op->kap32, &op->rsf64, &op->mf64, &op->rsa64,
                            &op->ma64, &op->rsf32, &op->mf32, &op->rsa32, &op->ma32, &op->V, &op->X, &op->Z, &op->Bh,
                            &op->dpart, &op->npart, &op->cpart, &op->rhs32};
    for (const DevBuf *bf : bufs) total += (int64_t)bf->bytes;
    if (op->mg)
        for (const auto &L : op->mg->levels)
            total += (int64_t)(L->rho.bytes + L->kap.bytes + L->dinv.bytes + L->bb.bytes + L->xx.bytes + L->tt.bytes);
    return total;
}

}  // extern "C"

// ---- Homogeneous surface condensation ------------------------------------

struct hz_surface {
    int nx = 0, ny = 0, nz = 0;
    int64_t nxy = 0;
    double dx = 0;
    DevBuf vx, ix, vy, iy, shift, zd, zu, zl, h, schur;
    DevBuf modes;                         // surface modes of the last solve_flux
    bool solved = false;
    int32_t ncols = 0;
    DevBuf node_ptr, node_col, node_val;  // masses grouped by plane node (CSR)
    DevBuf col_ptr, col_node, col_val;    // masses grouped by emitter (CSR of the transpose)
    DevBuf t1, t2, sel;                   // stage-1 products, planes / observation kernels, plane selectors
};

namespace {

void check_surface(const hz_surface *s) {
    if (!s) throw BadArgument("null surface handle");
}

// C[p] = A[p] B[p] for `count` products, in grid-z slices of at most 65535.
void zgemm(ZgemmArgs g, int count, hipStream_t st) {
    for (int p0 = 0; p0 < count; p0 += 65535) {
        const int c = std::min(65535, count - p0);
        ZgemmArgs q = g;
        q.a = g.a + p0 * g.a_sp;
        q.b = g.b + p0 * g.b_sp;
        q.c = g.c + p0 * g.c_sp;
        if (g.sel) q.sel = g.sel + p0;
        else if (g.w) q.w = g.w + p0 * g.w_sp;
        hipLaunchKernelGGL(k_zgemm, dim3((g.m + ZG_TILE - 1) / ZG_TILE, (g.n + ZG_TILE - 1) / ZG_TILE, c), dim3(256),
                           0, st, q);
        HZ_LAUNCHED();
    }
}

// Y[p] = op(L) (W[p] o X[p]) op(R)^T with L, R the x / y matrices of the
// forward (ix, iy) or inverse (vx, vy) transform; op = transpose when asked.
// W[p] = h[sel[p]] (no weight for sel[p] < 0) when sel is given; Y is
// divided by D elementwise when D is given.  Stage 1 goes through S->t1.
void modal_product(hz_surface *S, bool inverse, bool trans_l, bool trans_r, const cx<double> *x, int64_t x_sp,
                   const int *sel, const cx<double> *d, cx<double> *y, int count, hipStream_t st) {
    const int64_t nx = S->nx, ny = S->ny, nxy = S->nxy;
    S->t1.ensure((size_t)count * nxy * sizeof(cx<double>));
    const cx<double> *L = (inverse ? S->vx : S->ix).as<cx<double>>();
    const cx<double> *R = (inverse ? S->vy : S->iy).as<cx<double>>();
    ZgemmArgs g{};
    // Stage 1: T = (W o X) op(R)^T, nx x ny with inner ny.
    g.m = S->nx, g.n = S->ny, g.k = S->ny;
    g.a = x, g.a_si = 1, g.a_sk = nx, g.a_sp = x_sp;
    g.w = sel ? S->h.as<cx<double>>() : nullptr, g.w_sp = nxy, g.sel = sel;
    g.b = R, g.b_sp = 0;
    if (trans_r) g.b_sk = 1, g.b_sj = ny;
    else g.b_sk = ny, g.b_sj = 1;
    g.c = S->t1.as<cx<double>>(), g.ldc = nx, g.c_sp = nxy, g.d = nullptr;
    zgemm(g, count, st);
    // Stage 2: Y = op(L) T, nx x ny with inner nx.
    ZgemmArgs f{};
    f.m = S->nx, f.n = S->ny, f.k = S->nx;
    f.a = L, f.a_sp = 0;
    if (trans_l) f.a_si = nx, f.a_sk = 1;
    else f.a_si = 1, f.a_sk = nx;
    f.b = S->t1.as<cx<double>>(), f.b_sk = 1, f.b_sj = nx, f.b_sp = nxy;
    f.c = y, f.ldc = nx, f.c_sp = nxy, f.d = d;
    zgemm(f, count, st);
}

// Host corner indices checked against the grid; uploads low and frac next
// to each other in S->sel (int32 then double, 8-byte aligned).
void upload_points(hz_surface *S, const int32_t *low, const double *frac, int32_t n, int extra_ints,
                   int **d_low, double **d_frac, int **d_extra) {
    if (n < 1 || !low || !frac) throw BadArgument("need at least one point with low corners and fractions");
    for (int32_t p = 0; p < n; ++p) {
        const int lim[3] = {S->nx - 2, S->ny - 2, S->nz - 2};
        for (int a = 0; a < 3; ++a) {
            if (low[3 * p + a] < 0 || low[3 * p + a] > lim[a]) throw BadArgument("point corner outside the grid");
            if (!(frac[3 * p + a] >= 0.0 && frac[3 * p + a] <= 1.0)) throw BadArgument("point fraction outside [0, 1]");
        }
    }
    const size_t ints = (size_t)3 * n + (size_t)extra_ints;
    const size_t ioff = (ints * sizeof(int) + 7) / 8 * 8;
    S->sel.ensure(ioff + (size_t)3 * n * sizeof(double));
    char *base = S->sel.as<char>();
    HZ_HIP(hipMemcpy(base, low, (size_t)3 * n * sizeof(int), hipMemcpyHostToDevice));
    HZ_HIP(hipMemcpy(base + ioff, frac, (size_t)3 * n * sizeof(double), hipMemcpyHostToDevice));
    *d_low = reinterpret_cast<int *>(base);
    *d_frac = reinterpret_cast<double *>(base + ioff);
    if (d_extra) *d_extra = reinterpret_cast<int *>(base) + 3 * n;
}

}  // namespace

extern "C" {

hz_status hz_z_transfer(const void *shift, const void *diagonal, const void *upper, const void *lower, int32_t nz,
                        int64_t modes, void *h, void *schur, void *stream) {
    return guarded([&] {
        if (!shift || !diagonal || !upper || !lower || !h || !schur) throw BadArgument("null device pointer");
        if (nz < 2 || modes < 1) throw BadArgument("need nz >= 2 and at least one mode");
        hipLaunchKernelGGL(k_z_transfer, dim3((unsigned)((modes + 255) / 256)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), static_cast<const cx<double> *>(shift),
                           static_cast<const cx<double> *>(diagonal), static_cast<const cx<double> *>(upper),
                           static_cast<const cx<double> *>(lower), (int)nz, modes, static_cast<cx<double> *>(h),
                           static_cast<cx<double> *>(schur));
        HZ_LAUNCHED();
    });
}

hz_status hz_surface_create(const hz_surface_desc *d, hz_surface **out) {
    return guarded([&] {
        if (!d || !out) throw BadArgument("null descriptor or output");
        *out = nullptr;
        if (d->nx < 2 || d->ny < 2 || d->nz < 2) throw BadArgument("need nx, ny, nz >= 2");
        if (!(d->dx > 0) || !std::isfinite(d->dx)) throw BadArgument("dx must be finite and positive");
        if (!d->vx || !d->ix || !d->vy || !d->iy || !d->shift || !d->z_diagonal || !d->z_upper || !d->z_lower)
            throw BadArgument("missing surface table");
        auto S = std::make_unique<hz_surface>();
        S->nx = d->nx, S->ny = d->ny, S->nz = d->nz, S->dx = d->dx;
        S->nxy = (int64_t)d->nx * d->ny;
        upload(S->vx, d->vx, (size_t)2 * d->nx * d->nx);
        upload(S->ix, d->ix, (size_t)2 * d->nx * d->nx);
        upload(S->vy, d->vy, (size_t)2 * d->ny * d->ny);
        upload(S->iy, d->iy, (size_t)2 * d->ny * d->ny);
        upload(S->shift, d->shift, (size_t)2 * S->nxy);
        upload(S->zd, d->z_diagonal, (size_t)2 * d->nz);
        upload(S->zu, d->z_upper, (size_t)2 * (d->nz - 1));
        upload(S->zl, d->z_lower, (size_t)2 * (d->nz - 1));
        S->h.ensure((size_t)(d->nz - 1) * S->nxy * sizeof(cx<double>));
        S->schur.ensure((size_t)S->nxy * sizeof(cx<double>));
        hipLaunchKernelGGL(k_z_transfer, dim3((unsigned)((S->nxy + 255) / 256)), dim3(256), 0, nullptr,
                           S->shift.as<cx<double>>(), S->zd.as<cx<double>>(), S->zu.as<cx<double>>(),
                           S->zl.as<cx<double>>(), S->nz, S->nxy, S->h.as<cx<double>>(), S->schur.as<cx<double>>());
        HZ_LAUNCHED();
        HZ_HIP(hipDeviceSynchronize());
        *out = S.release();
    });
}

hz_status hz_surface_destroy(hz_surface *s) {
    return guarded([&] {
        if (s) {
            HZ_HIP(hipDeviceSynchronize());
            delete s;
        }
    });
}

hz_status hz_surface_tables(const hz_surface *s, void *h, void *schur, void *stream) {
    return guarded([&] {
        check_surface(s);
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (h)
            HZ_HIP(hipMemcpyAsync(h, s->h.p, (size_t)(s->nz - 1) * s->nxy * sizeof(cx<double>),
                                  hipMemcpyDeviceToDevice, st));
        if (schur) HZ_HIP(hipMemcpyAsync(schur, s->schur.p, (size_t)s->nxy * sizeof(cx<double>), hipMemcpyDeviceToDevice, st));
    });
}

hz_status hz_surface_set_sources(hz_surface *s, int32_t num_columns, int64_t nnz, const int64_t *node,
                                 const int32_t *column, const double *value) {
    return guarded([&] {
        check_surface(s);
        if (num_columns < 1 || nnz < 1 || !node || !column || !value) throw BadArgument("empty source list");
        for (int64_t e = 0; e < nnz; ++e)
            if (node[e] < 0 || node[e] >= s->nxy || column[e] < 0 || column[e] >= num_columns || !std::isfinite(value[e]))
                throw BadArgument("source entry out of range or not finite");
        // CSR by plane node and by emitter, each sorted by the other index.
        auto build = [&](const int64_t *row_of, int64_t nrows, const auto &col_of, DevBuf &ptr, DevBuf &col,
                         DevBuf &val) {
            std::vector<int64_t> order(nnz), rp(nrows + 1, 0);
            for (int64_t e = 0; e < nnz; ++e) order[e] = e;
            std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
                return row_of[a] != row_of[b] ? row_of[a] < row_of[b] : col_of(a) < col_of(b);
            });
            std::vector<int32_t> cc(nnz);
            std::vector<double> vv(nnz);
            for (int64_t t = 0; t < nnz; ++t) {
                const int64_t e = order[t];
                rp[row_of[e] + 1] += 1;
                cc[t] = (int32_t)col_of(e);
                vv[t] = value[e];
            }
            for (int64_t r = 0; r < nrows; ++r) rp[r + 1] += rp[r];
            upload(ptr, rp.data(), rp.size());
            upload(col, cc.data(), cc.size());
            upload(val, vv.data(), vv.size());
        };
        std::vector<int64_t> col64(column, column + nnz);
        build(node, s->nxy, [&](int64_t e) { return (int64_t)column[e]; }, s->node_ptr, s->node_col, s->node_val);
        build(col64.data(), num_columns, [&](int64_t e) { return node[e]; }, s->col_ptr, s->col_node, s->col_val);
        s->ncols = num_columns;
    });
}

hz_status hz_surface_transform(hz_surface *s, int32_t inverse, const void *x, void *y, int32_t count,
                               void *stream) {
    return guarded([&] {
        check_surface(s);
        if (!x || !y || count < 1) throw BadArgument("need device planes and count >= 1");
        if (x == y) throw BadArgument("x and y must not alias");
        modal_product(s, inverse != 0, false, false, static_cast<const cx<double> *>(x), s->nxy, nullptr, nullptr,
                      static_cast<cx<double> *>(y), count, static_cast<hipStream_t>(stream));
    });
}

hz_status hz_surface_solve_flux(hz_surface *s, const void *w, void *rhs, void *modes, void *surface, void *stream) {
    return guarded([&] {
        check_surface(s);
        if (s->ncols == 0) throw BadArgument("no source masses set");
        if (!w) throw BadArgument("w must be a device pointer");
        hipStream_t st = static_cast<hipStream_t>(stream);
        s->t2.ensure((size_t)s->nxy * sizeof(cx<double>));
        s->modes.ensure((size_t)s->nxy * sizeof(cx<double>));
        cx<double> *load = rhs ? static_cast<cx<double> *>(rhs) : s->t2.as<cx<double>>();
        const double scale = -4.0 * M_PI / (s->dx * s->dx * s->dx);
        const int64_t threads = s->nxy * WAVE;
        hipLaunchKernelGGL(k_csr_rows, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, (int)s->nxy, 1,
                           s->node_ptr.as<int64_t>(), s->node_col.as<int>(), s->node_val.as<double>(),
                           static_cast<const cx<double> *>(w), (int64_t)0, scale, load, s->nxy);
        HZ_LAUNCHED();
        modal_product(s, false, false, false, load, s->nxy, nullptr, s->schur.as<cx<double>>(), s->modes.as<cx<double>>(),
                      1, st);
        s->solved = true;
        if (modes)
            HZ_HIP(hipMemcpyAsync(modes, s->modes.p, (size_t)s->nxy * sizeof(cx<double>), hipMemcpyDeviceToDevice, st));
        if (surface)
            modal_product(s, true, false, false, s->modes.as<cx<double>>(), 0, nullptr, nullptr,
                          static_cast<cx<double> *>(surface), 1, st);
    });
}

hz_status hz_surface_planes(hz_surface *s, const int32_t *z, int32_t count, void *out, void *stream) {
    return guarded([&] {
        check_surface(s);
        if (!s->solved) throw BadArgument("no solved field: call hz_surface_solve_flux first");
        if (!z || !out || count < 1) throw BadArgument("need plane indices and an output");
        std::vector<int> sel(count);
        for (int p = 0; p < count; ++p) {
            if (z[p] < 0 || z[p] >= s->nz) throw BadArgument("z index outside grid");
            sel[p] = z[p] - 1;
        }
        upload(s->sel, sel.data(), sel.size());
        hipStream_t st = static_cast<hipStream_t>(stream);
        modal_product(s, true, false, false, s->modes.as<cx<double>>(), 0, s->sel.as<int>(), nullptr,
                      static_cast<cx<double> *>(out), count, st);
        HZ_HIP(hipStreamSynchronize(st));  // the selectors are reused by the next call
    });
}

hz_status hz_surface_sample(hz_surface *s, const int32_t *low, const double *frac, int32_t npoints, void *out,
                            void *stream) {
    return guarded([&] {
        check_surface(s);
        if (!s->solved) throw BadArgument("no solved field: call hz_surface_solve_flux first");
        if (!out) throw BadArgument("out must be a device pointer");
        hipStream_t st = static_cast<hipStream_t>(stream);
        // The distinct planes the points touch, ascending; two slots per point.
        if (npoints < 1 || !low) throw BadArgument("need at least one point with low corners and fractions");
        std::vector<int> planes;
        for (int p = 0; p < npoints; ++p)
            for (int dz = 0; dz < 2; ++dz) planes.push_back(low[3 * p + 2] + dz);
        std::sort(planes.begin(), planes.end());
        planes.erase(std::unique(planes.begin(), planes.end()), planes.end());
        std::vector<int> extra(planes.size() + 2 * (size_t)npoints);
        for (size_t q = 0; q < planes.size(); ++q) extra[q] = planes[q] - 1;
        for (int p = 0; p < npoints; ++p)
            for (int dz = 0; dz < 2; ++dz)
                extra[planes.size() + 2 * p + dz] =
                    (int)(std::lower_bound(planes.begin(), planes.end(), low[3 * p + 2] + dz) - planes.begin());
        int *d_low, *d_extra;
        double *d_frac;
        upload_points(s, low, frac, npoints, (int)extra.size(), &d_low, &d_frac, &d_extra);
        HZ_HIP(hipMemcpy(d_extra, extra.data(), extra.size() * sizeof(int), hipMemcpyHostToDevice));
        const int np = (int)planes.size();
        s->t2.ensure((size_t)np * s->nxy * sizeof(cx<double>));
        modal_product(s, true, false, false, s->modes.as<cx<double>>(), 0, d_extra, nullptr, s->t2.as<cx<double>>(), np,
                      st);
        hipLaunchKernelGGL(k_sample_planes, dim3((npoints + 255) / 256), dim3(256), 0, st, s->nx, s->ny, (int)npoints,
                           (const cx<double> *)s->t2.as<cx<double>>(), (const int *)(d_extra + np), (const int *)d_low,
                           (const double *)d_frac, static_cast<cx<double> *>(out));
        HZ_LAUNCHED();
        HZ_HIP(hipStreamSynchronize(st));  // the index arrays are reused by the next call
    });
}

hz_status hz_surface_target_response(hz_surface *s, const int32_t *low, const double *frac, int32_t npoints,
                                     void *out, void *stream) {
    return guarded([&] {
        check_surface(s);
        if (s->ncols == 0) throw BadArgument("no source masses set");
        if (!out) throw BadArgument("out must be a device pointer");
        hipStream_t st = static_cast<hipStream_t>(stream);
        int *d_low;
        double *d_frac;
        upload_points(s, low, frac, npoints, 0, &d_low, &d_frac, nullptr);
        const double scale = -4.0 * M_PI / (s->dx * s->dx * s->dx);
        constexpr int CHUNK = 128;  // observation kernels per pass
        for (int p0 = 0; p0 < npoints; p0 += CHUNK) {
            const int c = std::min(CHUNK, npoints - p0);
            s->t2.ensure((size_t)2 * c * s->nxy * sizeof(cx<double>));
            cx<double> *obs = s->t2.as<cx<double>>();
            cx<double> *ker = obs + (size_t)c * s->nxy;
            hipLaunchKernelGGL(k_observation, dim3((unsigned)((s->nxy + 255) / 256), c), dim3(256), 0, st, s->nx, s->ny,
                               s->nz, s->vx.as<cx<double>>(), s->vy.as<cx<double>>(), s->h.as<cx<double>>(),
                               s->schur.as<cx<double>>(), (const int *)(d_low + 3 * p0),
                               (const double *)(d_frac + 3 * p0), obs);
            HZ_LAUNCHED();
            // kernel_p = ix^T O_p iy
            modal_product(s, false, true, true, obs, s->nxy, nullptr, nullptr, ker, c, st);
            const int64_t threads = (int64_t)s->ncols * c * WAVE;
            hipLaunchKernelGGL(k_csr_rows, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, (int)s->ncols, c,
                               s->col_ptr.as<int64_t>(), s->col_node.as<int>(), s->col_val.as<double>(),
                               (const cx<double> *)ker, s->nxy, scale,
                               static_cast<cx<double> *>(out) + (int64_t)p0 * s->ncols, (int64_t)s->ncols);
            HZ_LAUNCHED();
        }
        HZ_HIP(hipStreamSynchronize(st));  // the point arrays are reused by the next call
    });
}

}  // extern "C"

// ---- Source deposition ------------------------------------------------------

extern "C" {

hz_status hz_deposit_sources(int32_t m, const double *centers, int32_t q, const double *offsets,
                             const double *weights, double dx, const int32_t *boxes, int32_t capacity, double *mass) {
    return guarded([&] {
        if (m < 1 || q < 1 || capacity < 1) throw BadArgument("need m, q, capacity >= 1");
        if (!centers || !offsets || !weights || !boxes || !mass) throw BadArgument("null host array");
        if (!(dx > 0) || !std::isfinite(dx)) throw BadArgument("dx must be finite and positive");
        for (int e = 0; e < m; ++e) {
            const int32_t *b = boxes + 4 * e;
            if (b[2] < 1 || b[3] < 1 || (int64_t)b[2] * b[3] > capacity || b[0] < 0 || b[1] < 0)
                throw BadArgument("emitter box empty, negative or larger than capacity");
        }
        DevBuf dc, doff, dw, db, dm;
        upload(dc, centers, (size_t)2 * m);
        upload(doff, offsets, (size_t)2 * q);
        upload(dw, weights, (size_t)q);
        upload(db, boxes, (size_t)4 * m);
        dm.ensure((size_t)m * capacity * sizeof(double));
        hipLaunchKernelGGL(k_deposit, dim3(m), dim3(DEPOSIT_CHUNK), 0, nullptr, dx, (int)q, dc.as<double>(),
                           doff.as<double>(), dw.as<double>(), db.as<int>(), (int)capacity, dm.as<double>());
        HZ_LAUNCHED();
        HZ_HIP(hipMemcpy(mass, dm.p, (size_t)m * capacity * sizeof(double), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

// ---- Scene rasterisation ------------------------------------------------------

extern "C" {

hz_status hz_rasterize_materials(int32_t nx, int32_t ny, int32_t nz, double dx, int32_t nprim,
                                 const hz_primitive *prims, double width, double rho_obstacle, double c_obstacle,
                                 double rho_medium, double c_medium, const double *rho_field, const double *c_field,
                                 double *rho, double *c, double *sdf, void *stream) {
    return guarded([&] {
        if (nx < 1 || ny < 1 || nz < 1) throw BadArgument("empty grid");
        if (!(dx > 0) || !std::isfinite(dx) || !(width >= 0) || !std::isfinite(width))
            throw BadArgument("dx must be positive and width non-negative");
        if (nprim < 0 || nprim > MAX_PRIMITIVES) throw BadArgument("between 0 and 16 primitives");
        if (nprim > 0 && !prims) throw BadArgument("null primitive list");
        if (!rho && !c && !sdf) throw BadArgument("no output requested");
        if (!(rho_obstacle > 0) || !(c_obstacle > 0) || (!rho_field && !(rho_medium > 0)) || (!c_field && !(c_medium > 0)))
            throw BadArgument("densities and sound speeds must be positive");
        MaterialArgs a{};
        a.nx = nx, a.ny = ny, a.nz = nz, a.dx = dx, a.width = width, a.nprim = nprim;
        for (int q = 0; q < nprim; ++q) {
            const hz_primitive &p = prims[q];
            if (p.kind < PRIM_SPHERE || p.kind > PRIM_ELLIPSOID) throw BadArgument("unknown primitive kind");
            const int need = p.kind == PRIM_SPHERE ? 1 : 3;
            for (int e = 0; e < 3; ++e) {
                if (!std::isfinite(p.center[e])) throw BadArgument("primitive centre not finite");
                if (e < need && !(p.size[e] > 0 && std::isfinite(p.size[e]))) throw BadArgument("primitive size must be positive");
                a.prim[q].c[e] = p.center[e];
                a.prim[q].s[e] = p.size[e];
            }
            a.prim[q].kind = p.kind;
        }
        a.rho_obs = rho_obstacle, a.c_obs = c_obstacle, a.rho_med = rho_medium, a.c_med = c_medium;
        a.rho_field = rho_field, a.c_field = c_field, a.rho = rho, a.c = c, a.sdf = sdf;
        const int64_t n = (int64_t)nx * ny * nz;
        hipLaunchKernelGGL(k_materials, dim3((unsigned)blocks_for(n, 256, 65536)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), a);
        HZ_LAUNCHED();
    });
}

}  // extern "C"

extern "C" {

hz_status hz_gaussian_filter(int32_t nx, int32_t ny, int32_t nz, int32_t radius, const double *weights, double *x,
                             void *stream) {
    return guarded([&] {
        if (nx < 1 || ny < 1 || nz < 1 || radius < 0) throw BadArgument("empty grid or negative radius");
        if (!weights || !x) throw BadArgument("null weights or grid");
        for (int j = 0; j < radius; ++j)
            if (weights[j] != weights[2 * radius - j]) throw BadArgument("the weights must be symmetric");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int64_t n = (int64_t)nx * ny * nz;
        DevBuf w, tmp;
        upload(w, weights, (size_t)2 * radius + 1);
        tmp.ensure((size_t)n * sizeof(double));
        double *src = x, *dst = tmp.as<double>();
        for (int axis = 0; axis < 3; ++axis) {
            hipLaunchKernelGGL(k_wrap_filter, dim3((unsigned)blocks_for(n, 256, 65536)), dim3(256), 0, st, nx, ny, nz,
                               axis, (int)radius, (const double *)w.as<double>(), (const double *)src, dst);
            HZ_LAUNCHED();
            std::swap(src, dst);
        }
        // Three passes leave the result in the scratch buffer.
        HZ_HIP(hipMemcpyAsync(x, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HZ_HIP(hipStreamSynchronize(st));  // the scratch buffer is freed on return
    });
}

}  // extern "C"
