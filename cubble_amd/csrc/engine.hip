// engine.hip -- the internal host runtime (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cstring>

#include "host_init.hpp"

namespace cubble_amd {

namespace {

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

#define E_TRY(call)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(e_ == hipErrorOutOfMemory ? CUBBLE_E_NOMEM : CUBBLE_E_HIP,          \
                        std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// Assigns every per-foam buffer (except the neighbour lists) inside one
// block starting at `base`; with base == nullptr only measures the size.
struct Carver {
    char *base;
    size_t off = 0;
    template <typename T>
    T *take(size_t count) {
        const size_t o = off;
        off = align256(off + std::max<size_t>(count, 1) * sizeof(T));
        return base ? reinterpret_cast<T *>(base + o) : nullptr;
    }
};

size_t carve_foam(Foam &F, char *base) {
    Carver L{base};
    const size_t cap = (size_t)F.cap;
    const int D = F.c.dim;
    double *xb[2][4] = {};
    void *vb[4][4] = {};
    // velocities, saved positions, error and overlap: FP32 in mixed precision
    const bool fvel = CUBBLE_PREC_MIXED == F.precision;
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 4; ++q)
            if (!(2 == q && 2 == D)) xb[a][q] = L.take<double>(cap);
    for (int a = 0; a < 4; ++a)
        for (int q = 0; q < 4; ++q)
            if (!(2 == q && 2 == D))
                vb[a][q] = fvel ? (void *)L.take<float>(cap) : (void *)L.take<double>(cap);
    for (int q = 0; q < 4; ++q) {
        F.r.x[q] = xb[0][q];
        F.r.xp[q] = xb[1][q];
        F.r.v[q] = vb[0][q];
        F.r.vp[q] = vb[1][q];
        F.r.vo[q] = vb[2][q];
        F.r.vs[q] = vb[3][q];
        F.r.saved[q] = (2 == q && 2 == D) ? nullptr
                       : fvel ? (void *)L.take<float>(cap) : (void *)L.take<double>(cap);
    }
    F.r.path = L.take<double>(cap);
    F.r.path_new = L.take<double>(cap);
    F.r.err = fvel ? (void *)L.take<float>(cap) : (void *)L.take<double>(cap);
    F.r.err_alt = fvel ? (void *)L.take<float>(cap) : (void *)L.take<double>(cap);
    for (int q = 0; q < 3; ++q) {
        F.r.wc[q] = q < D ? L.take<int32_t>(cap) : nullptr;
        F.r.wc_alt[q] = q < D ? L.take<int32_t>(cap) : nullptr;
    }
    F.r.idx = L.take<int32_t>(cap);
    F.r.idx_alt = L.take<int32_t>(cap);
    Scratch &S = F.s;
    // FP64: the 32-byte predicted record; mixed precision: the 16-byte one
    // (k_correct recomputes the FP64 prediction)
    const bool mixed = CUBBLE_PREC_MIXED == F.precision;
    S.p4 = mixed ? nullptr : L.take<D4>(cap);
    S.q4 = mixed ? L.take<Q4>(cap) : nullptr;
    S.x4 = L.take<D4>(cap);
    S.ov = fvel ? (void *)L.take<float>(cap) : (void *)L.take<double>(cap);
    S.del = L.take<uint8_t>(cap);
    S.dirty = L.take<uint8_t>(cap);
    S.nn = L.take<int32_t>(cap);
    S.deg = L.take<int32_t>(cap);
    S.cell_of = L.take<int32_t>(cap);
    S.members = L.take<int32_t>(cap);
    S.del_list = L.take<int32_t>(cap);
    S.plan_src = L.take<int32_t>(cap);
    S.moved_to = L.take<int32_t>(cap);
    S.cell_count = L.take<int32_t>((size_t)F.cell_cap);
    S.cell_cursor = L.take<int32_t>((size_t)F.cell_cap);
    S.cell_start = L.take<int32_t>((size_t)F.cell_cap + 1);
    // block partials sized for the smallest launch shapes (one bubble per
    // thread, kMinSearchTile bubbles per search tile), so retuning never
    // reallocates
    const size_t nblk_max = ((size_t)F.cap + kBlock - 1) / kBlock;
    const size_t sblk_max = (size_t)F.cap / kMinSearchTile + 1;
    S.part = L.take<double>(nblk_max * kPartD);
    S.part_i = L.take<int32_t>(nblk_max * kPartI);
    S.cpart = L.take<int32_t>((size_t)F.cblk);
    S.cpart_i = L.take<int32_t>(sblk_max * kPartI);
    return L.off;
}

// Launch shapes: bubble-space blocks and neighbour-search tiles.
void shape_foam(Foam &F, const Tuning &t) {
    const int cap = F.cap;
    F.ipt = t.ipt > 0 ? t.ipt : std::max(1, (cap + kBlock * kMaxBlocks - 1) / (kBlock * kMaxBlocks));
    F.nblk = std::max(1, (cap + kBlock * F.ipt - 1) / (kBlock * F.ipt));
    F.stile = std::max(kMinSearchTile, t.stile > 0 ? t.stile : 256);
    // tile t holds the cells whose first bubble is in [t B, t B + B); bubbles
    // are below the capacity, so t runs to cap / B
    F.sblk = F.cap / F.stile + 1;
}

void size_foam(Foam &F, int cap, int nbr_cap, int cell_cap, const Tuning &t) {
    F.cap = cap;
    F.nbr_cap = nbr_cap;
    F.cell_cap = cell_cap;
    F.cipt = std::max(1, (cell_cap + kBlock * 1024 - 1) / (kBlock * 1024));
    F.cblk = std::max(1, (cell_cap + kBlock * F.cipt - 1) / (kBlock * F.cipt));
    shape_foam(F, t);
}

int default_nbr_cap(int dim) { return 3 == dim ? 48 : 24; }

}  // namespace

// ---------------------------------------------------------------- arena
DeviceArena::~DeviceArena() {
    for (auto &c : chunks_) (void)hipFree(c.base);
}

bool DeviceArena::add_chunk(size_t bytes) {
    bytes = align256(bytes);
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return false;
    Chunk c;
    c.base = static_cast<char *>(p);
    c.size = bytes;
    c.free[0] = bytes;
    chunks_.push_back(std::move(c));
    capacity_ += bytes;
    return true;
}

int DeviceArena::init(size_t first_chunk, bool growable, std::string &err) {
    growable_ = growable;
    if (first_chunk > 0 && !add_chunk(first_chunk)) {
        err = "hipMalloc of the device arena failed";
        return CUBBLE_E_NOMEM;
    }
    return CUBBLE_OK;
}

void *DeviceArena::alloc(size_t bytes) {
    bytes = align256(std::max<size_t>(bytes, 256));
    for (int pass = 0; pass < 2; ++pass) {
        for (auto &c : chunks_) {
            for (auto it = c.free.begin(); it != c.free.end(); ++it) {
                if (it->second < bytes) continue;
                const size_t off = it->first, len = it->second;
                c.free.erase(it);
                if (len > bytes) c.free[off + bytes] = len - bytes;
                void *p = c.base + off;
                live_[p] = bytes;
                in_use_ += bytes;
                return p;
            }
        }
        if (pass == 0) {
            if (!growable_) return nullptr;
            if (!add_chunk(std::max(bytes, (size_t)64 << 20))) return nullptr;
        }
    }
    return nullptr;
}

void DeviceArena::release(void *p) {
    auto it = live_.find(p);
    if (it == live_.end()) return;
    size_t len = it->second;
    live_.erase(it);
    in_use_ -= len;
    for (auto &c : chunks_) {
        char *q = static_cast<char *>(p);
        if (q < c.base || q >= c.base + c.size) continue;
        size_t off = (size_t)(q - c.base);
        auto nx = c.free.lower_bound(off);
        if (nx != c.free.end() && nx->first == off + len) {
            len += nx->second;
            nx = c.free.erase(nx);
        }
        if (nx != c.free.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                pv->second += len;
                return;
            }
        }
        c.free[off] = len;
        return;
    }
}

size_t DeviceArena::largest_free() const {
    size_t m = 0;
    for (const auto &c : chunks_)
        for (const auto &f : c.free) m = std::max(m, f.second);
    return m;
}

size_t foam_footprint(int dim, int precision, int cap, int nbr_cap, int cell_cap,
                      const Tuning &t) {
    Foam F;
    std::memset(&F, 0, sizeof(F));
    F.c.dim = dim;
    F.precision = precision;
    size_foam(F, cap, nbr_cap > 0 ? nbr_cap : default_nbr_cap(dim), cell_cap, t);
    return carve_foam(F, nullptr) + align256((size_t)F.nbr_cap * cap * sizeof(int32_t));
}

// ---------------------------------------------------------------- engine
Engine::~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (void *p : {(void *)dfoams_, (void *)bt_bub_, (void *)bt_cell_, (void *)bt_scell_})
        if (p) (void)hipFree(p);
    if (pinned_) (void)hipHostFree(pinned_);
    if (stream) (void)hipStreamDestroy(stream);
}

int Engine::init(size_t arena_bytes, bool growable) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CUBBLE_E_HIP, "no HIP device found");
    if (hipGetDevice(&device) != hipSuccess) device = 0;
    E_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    std::string e;
    const int rc = arena.init(arena_bytes, growable, e);
    if (rc) return fail(rc, e);
    return CUBBLE_OK;
}

int Engine::add_foam(const cubble_config &cfg, const Tuning &t, int *index) {
    std::vector<double> x4;
    InitOut io;
    std::string e;
    const int rc = generate_foam(cfg, x4, io, e);
    if (rc) return fail(rc, e);
    return add_generated(cfg, t, x4, io, index);
}

int Engine::add_generated(const cubble_config &cfg, const Tuning &t, const std::vector<double> &x4,
                          const InitOut &io, int *index) {
    if (!foams.empty() && cfg.precision != precision)
        return fail(CUBBLE_E_INVALID, "all foams of one batch must share the precision mode");
    precision = cfg.precision;
    const int n = cfg.num_start;
    FoamRec r;
    Foam &F = r.h;
    std::memset(&F, 0, sizeof(F));
    F.c = io.c;
    F.precision = cfg.precision;
    size_foam(F, n, cfg.nbr_cap > 0 ? cfg.nbr_cap : default_nbr_cap(cfg.dim), io.cell_cap, t);
    const size_t bytes = carve_foam(F, nullptr);
    r.block = arena.alloc(bytes);
    r.nbr_block = r.block ? arena.alloc((size_t)F.nbr_cap * n * sizeof(int32_t)) : nullptr;
    if (!r.block || !r.nbr_block) {
        if (r.block) arena.release(r.block);
        return fail(CUBBLE_E_NOMEM, "device memory for the foam is not available");
    }
    carve_foam(F, static_cast<char *>(r.block));
    F.s.nbr = static_cast<int32_t *>(r.nbr_block);
    E_TRY(hipMemsetAsync(r.block, 0, bytes, stream));
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == cfg.dim) continue;
        E_TRY(hipMemcpyAsync(F.r.x[q], x4.data() + (size_t)q * n, sizeof(double) * n,
                             hipMemcpyHostToDevice, stream));
    }
    std::vector<int32_t> iota(n);
    for (int i = 0; i < n; ++i) iota[i] = i;
    E_TRY(hipMemcpyAsync(F.r.idx, iota.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream));
    E_TRY(hipStreamSynchronize(stream));
    Dyn &d = F.d;
    d.count = n;
    d.dt = cfg.dt0;
    d.max_r_cell = io.max_r;
    d.need_search = 1;
    d.ts_min = kTsMinEmpty;
    d.ts_max = kTsMaxEmpty;
    r.num_end = cfg.num_end;
    r.max_steps = cfg.max_steps;
    tuning_ = t;
    foams.push_back(r);
    dev_index_.push_back(-1);
    if (index) *index = (int)foams.size() - 1;
    return CUBBLE_OK;
}

int Engine::drop_foam(int k) {
    if (k < 0 || k >= (int)foams.size()) return fail(CUBBLE_E_INVALID, "no such foam");
    // keep the device state of the remaining foams before the array changes
    const int rc = pull_headers();
    if (rc) return rc;
    for (auto &di : dev_index_) di = -2;  // host mirrors are current
    arena.release(foams[k].block);
    arena.release(foams[k].nbr_block);
    foams.erase(foams.begin() + k);
    dev_index_.erase(dev_index_.begin() + k);
    return CUBBLE_OK;
}

int Engine::pull_headers() {
    if (!dfoams_) return CUBBLE_OK;
    E_TRY(hipStreamSynchronize(stream));
    for (size_t i = 0; i < foams.size(); ++i)
        if (dev_index_[i] >= 0)
            E_TRY(hipMemcpy(&foams[i].h, dfoams_ + dev_index_[i], sizeof(Foam),
                            hipMemcpyDeviceToHost));
    return CUBBLE_OK;
}

int Engine::commit_layout() {
    E_TRY(hipStreamSynchronize(stream));
    const int nf = (int)foams.size();
    // device state of foams already on the device is the truth
    int rc = pull_headers();
    if (rc) return rc;
    const Foam *old_foams = dfoams_;
    const int2 *old_bt[3] = {bt_bub_, bt_cell_, bt_scell_};
    const int old_nb[3] = {nb_bub_, nb_cell_, nb_scell_};
    const int old_nf = committed_nf_;
    if (nf > dfoams_cap_) {
        if (dfoams_) (void)hipFree(dfoams_);
        dfoams_ = nullptr;
        const int cap = std::max(nf, 2 * dfoams_cap_);
        E_TRY(hipMalloc(&dfoams_, sizeof(Foam) * std::max(cap, 1)));
        dfoams_cap_ = cap;
        if (pinned_) (void)hipHostFree(pinned_);
        pinned_ = nullptr;
        E_TRY(hipHostMalloc((void **)&pinned_, sizeof(Dyn) * std::max(cap, 1), hipHostMallocDefault));
        pinned_cap_ = cap;
    }
    std::vector<Foam> hs(std::max(nf, 1));
    std::vector<int2> bb, bc, bs;
    for (int f = 0; f < nf; ++f) {
        hs[f] = foams[f].h;
        for (int l = 0; l < hs[f].nblk; ++l) bb.push_back(make_int2(f, l));
        for (int l = 0; l < hs[f].cblk; ++l) bc.push_back(make_int2(f, l));
        for (int l = 0; l < hs[f].sblk; ++l) bs.push_back(make_int2(f, l));
        dev_index_[f] = f;
    }
    if (nf > 0) E_TRY(hipMemcpy(dfoams_, hs.data(), sizeof(Foam) * nf, hipMemcpyHostToDevice));
    auto upload = [&](int2 *&buf, size_t &cap, const std::vector<int2> &v) -> int {
        if (v.size() > cap) {
            if (buf) (void)hipFree(buf);
            buf = nullptr;
            cap = std::max(v.size(), 2 * cap);
            E_TRY(hipMalloc(&buf, sizeof(int2) * cap));
        }
        if (!v.empty()) E_TRY(hipMemcpy(buf, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice));
        return CUBBLE_OK;
    };
    if ((rc = upload(bt_bub_, bt_bub_cap_, bb)) || (rc = upload(bt_cell_, bt_cell_cap_, bc)) ||
        (rc = upload(bt_scell_, bt_scell_cap_, bs)))
        return rc;
    nb_bub_ = (int)bb.size();
    nb_cell_ = (int)bc.size();
    nb_scell_ = (int)bs.size();
    // the captured graph holds the header array, the block tables and the
    // grid sizes; it stays valid while none of them moved or changed size
    const bool same = old_foams == dfoams_ && old_bt[0] == bt_bub_ && old_bt[1] == bt_cell_ &&
                      old_bt[2] == bt_scell_ && old_nb[0] == nb_bub_ && old_nb[1] == nb_cell_ &&
                      old_nb[2] == nb_scell_ && old_nf == nf;
    committed_nf_ = nf;
    if (graph_ && !same) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    return CUBBLE_OK;
}

int Engine::retune(const Tuning &t) {
    int rc = pull_headers();
    if (rc) return rc;
    for (auto &di : dev_index_) di = -2;  // host mirrors are current
    tuning_ = t;
    for (auto &r : foams) shape_foam(r.h, t);
    return commit_layout();
}

void Engine::free_snapshot(Snapshot &s) {
    for (void *p : s.blk)
        if (p) (void)hipFree(p);
    for (void *p : s.nbr)
        if (p) (void)hipFree(p);
    s = Snapshot();
}

int Engine::snapshot(Snapshot &s) {
    free_snapshot(s);
    int rc = pull_headers();
    if (rc) return rc;
    for (const FoamRec &r : foams) {
        Foam tmp = r.h;
        const size_t bb = carve_foam(tmp, nullptr);
        const size_t nb = (size_t)r.h.nbr_cap * r.h.cap * sizeof(int32_t);
        void *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, bb) != hipSuccess || hipMalloc(&b, nb) != hipSuccess) {
            if (a) (void)hipFree(a);
            free_snapshot(s);
            return fail(CUBBLE_E_NOMEM, "no device memory for a state snapshot");
        }
        s.h.push_back(r.h);
        s.blk.push_back(a);
        s.nbr.push_back(b);
        s.blk_bytes.push_back(bb);
        s.nbr_bytes.push_back(nb);
        E_TRY(hipMemcpyAsync(a, r.block, bb, hipMemcpyDeviceToDevice, stream));
        E_TRY(hipMemcpyAsync(b, r.nbr_block, nb, hipMemcpyDeviceToDevice, stream));
    }
    E_TRY(hipStreamSynchronize(stream));
    return CUBBLE_OK;
}

int Engine::restore(const Snapshot &s) {
    if (s.h.size() != foams.size()) return fail(CUBBLE_E_STATE, "snapshot of other foams");
    int rc = pull_headers();
    if (rc) return rc;
    for (size_t k = 0; k < foams.size(); ++k) {
        FoamRec &r = foams[k];
        // lists only grow, and the first columns of the grown block hold the
        // saved ones (column q of bubble i is nbr[q * cap + i])
        if (s.h[k].nbr_cap > r.h.nbr_cap || s.h[k].cap != r.h.cap)
            return fail(CUBBLE_E_STATE, "snapshot does not fit the foam");
        E_TRY(hipMemcpyAsync(r.block, s.blk[k], s.blk_bytes[k], hipMemcpyDeviceToDevice, stream));
        E_TRY(hipMemcpyAsync(r.nbr_block, s.nbr[k], s.nbr_bytes[k], hipMemcpyDeviceToDevice,
                             stream));
        Foam f = s.h[k];
        f.s.nbr = r.h.s.nbr;
        f.nbr_cap = r.h.nbr_cap;
        f.ipt = r.h.ipt;
        f.nblk = r.h.nblk;
        f.stile = r.h.stile;
        f.sblk = r.h.sblk;
        r.h = f;
    }
    E_TRY(hipStreamSynchronize(stream));
    for (auto &di : dev_index_) di = -2;
    return commit_layout();
}

Grid Engine::grid() const {
    Grid g;
    g.foams = dfoams_;
    g.bt_bub = bt_bub_;
    g.nb_bub = nb_bub_;
    g.bt_cell = bt_cell_;
    g.nb_cell = nb_cell_;
    g.bt_scell = bt_scell_;
    g.nb_scell = nb_scell_;
    g.nfoams = (int)foams.size();
    g.stream = stream;
    g.walls = false;
    for (const FoamRec &r : foams)
        for (int a = 0; a < r.h.c.dim; ++a) g.walls = g.walls || r.h.c.wall[a] != 0;
    return g;
}

void Engine::enqueue_attempt(uint32_t flags) {
    const Grid g = grid();
    launch_predict(g, flags);
    if (CUBBLE_PREC_MIXED == precision) launch_pair_mixed(g, flags);
    else launch_pair(g, flags);
    launch_sums(g);
    launch_correct(g, flags, CUBBLE_PREC_MIXED == precision);
}

void Engine::enqueue_group(uint32_t flags) {
    const Grid g = grid();
    enqueue_attempt(flags);
    launch_remove(g);
    launch_search(g);
}

int Engine::capture(uint32_t flags, int groups) {
    // the graph holds the kernel variants chosen for the foams it was
    // captured with (walls or not)
    const bool walls = grid().walls;
    if (graph_ && graph_flags_ == flags && graph_groups_ == groups && graph_walls_ == walls)
        return CUBBLE_OK;
    if (graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    hipGraph_t gr = nullptr;
    E_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < groups; ++k) enqueue_group(flags);
    E_TRY(hipStreamEndCapture(stream, &gr));
    const hipError_t e = hipGraphInstantiate(&graph_, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
        graph_ = nullptr;
        return fail(CUBBLE_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    graph_flags_ = flags;
    graph_groups_ = groups;
    graph_walls_ = walls;
    return CUBBLE_OK;
}

int Engine::sync() {
    E_TRY(hipStreamSynchronize(stream));
    return CUBBLE_OK;
}

int Engine::read_dyn_all(std::vector<Dyn> &out) {
    const int nf = (int)foams.size();
    out.resize(nf);
    if (nf == 0) return CUBBLE_OK;
    E_TRY(hipMemcpy2DAsync(pinned_, sizeof(Dyn), &dfoams_->d, sizeof(Foam), sizeof(Dyn), nf,
                           hipMemcpyDeviceToHost, stream));
    E_TRY(hipStreamSynchronize(stream));
    std::memcpy(out.data(), pinned_, sizeof(Dyn) * nf);
    return CUBBLE_OK;
}

int Engine::write_dyn(int k, const Dyn &d) {
    E_TRY(hipStreamSynchronize(stream));
    E_TRY(hipMemcpy(&dfoams_[k].d, &d, sizeof(Dyn), hipMemcpyHostToDevice));
    return CUBBLE_OK;
}

int Engine::read_header(int k, Foam &f) {
    E_TRY(hipStreamSynchronize(stream));
    E_TRY(hipMemcpy(&f, dfoams_ + k, sizeof(Foam), hipMemcpyDeviceToHost));
    return CUBBLE_OK;
}

int Engine::grow_lists(int k, int need, bool rebuild, bool keep) {
    Foam f;
    int rc = read_header(k, f);  // synchronises: no kernel uses the lists now
    if (rc) return rc;
    FoamRec &r = foams[k];
    const int ncap = std::max(need + need / 4 + 4, f.nbr_cap * 2);
    const size_t old_bytes = (size_t)f.nbr_cap * f.cap * sizeof(int32_t);
    void *nb = nullptr;
    if (keep) {
        nb = arena.alloc((size_t)ncap * f.cap * sizeof(int32_t));
        if (!nb) return fail(CUBBLE_E_NOMEM, "no device memory to grow the neighbour lists");
        // column q of bubble i is nbr[q * cap + i]: the old lists are the
        // first nbr_cap columns of the grown block
        E_TRY(hipMemcpyAsync(nb, r.nbr_block, old_bytes, hipMemcpyDeviceToDevice, stream));
        arena.release(r.nbr_block);
    } else {
        // the contents are rebuilt or replaced: free first, so that only the
        // new block has to fit
        arena.release(r.nbr_block);
        nb = arena.alloc((size_t)ncap * f.cap * sizeof(int32_t));
        if (!nb) {
            // the old block's room is free again: take a block of the old
            // size back (first fit: maybe elsewhere), so the foam stays whole
            // (halted by its overflow) for the caller to decide
            r.nbr_block = arena.alloc(old_bytes);
            int32_t *p = static_cast<int32_t *>(r.nbr_block);
            r.h.s.nbr = p;
            E_TRY(hipMemcpy(&dfoams_[k].s.nbr, &p, sizeof(p), hipMemcpyHostToDevice));
            return fail(CUBBLE_E_NOMEM, "no device memory to grow the neighbour lists");
        }
    }
    r.nbr_block = nb;
    int32_t *p = static_cast<int32_t *>(nb);
    r.h.s.nbr = p;
    r.h.nbr_cap = ncap;
    E_TRY(hipMemcpy(&dfoams_[k].s.nbr, &p, sizeof(p), hipMemcpyHostToDevice));
    E_TRY(hipMemcpy(&dfoams_[k].nbr_cap, &ncap, sizeof(ncap), hipMemcpyHostToDevice));
    if (!rebuild) return CUBBLE_OK;
    Dyn d = f.d;
    d.need_search = 1;
    d.halt = 0;
    d.overflow = 0;
    if ((rc = write_dyn(k, d))) return rc;
    launch_search_lists(grid());  // only foam k has its search flag set
    E_TRY(hipGetLastError());
    if ((rc = read_header(k, f))) return rc;
    if (f.d.overflow) return fail(CUBBLE_E_CAPACITY, "neighbour list capacity growth failed");
    f.d.searches -= 1;  // the rebuild is not a new search
    return write_dyn(k, f.d);
}

int Engine::fail_foam(int k, int code) {
    foams[k].failed = code;
    Foam f;
    int rc = read_header(k, f);
    if (rc) return rc;
    f.d.failed = 1;
    f.d.halt = 1;
    f.d.overflow = 0;
    f.d.need_search = 0;
    f.d.need_remove = 0;
    err.clear();
    return write_dyn(k, f.d);
}

int Engine::settle(std::vector<Dyn> &d) {
    int rc = read_dyn_all(d);
    if (rc) return rc;
    bool grew = false;
    for (int k = 0; k < (int)d.size(); ++k) {
        if (!d[k].overflow || foams[k].failed) continue;
        // the step was committed; other foams of the batch are unaffected
        rc = grow_lists(k, d[k].max_nn, true);
        if (CUBBLE_E_NOMEM == rc && isolate_failures) {
            if ((rc = fail_foam(k, CUBBLE_E_NOMEM))) return rc;
            grew = true;
            continue;
        }
        if (rc) return rc;
        Foam f;
        if ((rc = read_header(k, f))) return rc;
        f.d.halt = (stop_reached(f.d) || (f.d.count_stop > 0 && f.d.count <= f.d.count_stop)) ? 1 : 0;
        if ((rc = write_dyn(k, f.d))) return rc;
        grew = true;
    }
    // grow ahead of the next search when the last one came close to the
    // capacity: an overflow halts the foam until the host settles
    for (int k = 0; k < (int)d.size(); ++k) {
        const int cap = foams[k].h.nbr_cap;
        if (d[k].overflow || foams[k].failed || d[k].max_nn + std::max(4, cap / 8) <= cap)
            continue;
        rc = grow_lists(k, d[k].max_nn, false, true);
        // optional: without the memory the foam grows when it overflows
        if (CUBBLE_E_NOMEM == rc) {
            err.clear();
            continue;
        }
        if (rc) return rc;
    }
    if (grew) rc = read_dyn_all(d);
    return rc;
}

namespace {
// A foam is finished for the current run once the device halted it (step
// limit reached or dissolution stop) and no capacity rebuild is pending;
// returns the most steps any unfinished foam still has to go.
int64_t steps_left(const std::vector<Dyn> &d) {
    int64_t left = 0;
    for (const Dyn &x : d)
        if (!(x.halt && !x.overflow)) left = std::max<int64_t>(left, std::max<int64_t>(1, x.step_limit - x.steps));
    return left;
}
}  // namespace

int Engine::set_limits(int64_t steps, const std::vector<int64_t> *per_foam) {
    launch_set_limits(grid(), steps);
    E_TRY(hipGetLastError());
    if (!per_foam) return CUBBLE_OK;
    std::vector<Dyn> d;
    int rc = read_dyn_all(d);
    if (rc) return rc;
    for (size_t k = 0; k < d.size() && k < per_foam->size(); ++k) {
        const int64_t lim = d[k].steps + std::max<int64_t>(1, (*per_foam)[k]);
        E_TRY(hipMemcpy(&dfoams_[k].d.step_limit, &lim, sizeof(lim), hipMemcpyHostToDevice));
    }
    return CUBBLE_OK;
}

int Engine::integrate(uint32_t flags, std::vector<Dyn> &d) {
    int rc = set_limits(1, nullptr);
    if (rc) return rc;
    for (int tries = 0; tries < 100000; ++tries) {
        enqueue_group(flags);
        E_TRY(hipGetLastError());
        if ((rc = settle(d))) return rc;
        if (steps_left(d) == 0) {
            if ((rc = set_limits(0, nullptr))) return rc;
            return read_dyn_all(d);
        }
    }
    return fail(CUBBLE_E_STATE, "integrate: no accepted step after 100000 attempts");
}

int Engine::run_steps(uint32_t flags, int64_t steps, std::vector<Dyn> &d,
                      const std::vector<int64_t> *per_foam) {
    const int groups = std::max(1, tuning_.groups);
    if (steps <= 0) {
        // no steps: only prepare the device loop (capture its graph)
        const int rc = foams.empty() ? CUBBLE_OK : capture(flags, groups);
        return rc ? rc : read_dyn_all(d);
    }
    int rc = capture(flags, groups);
    if (rc) return rc;
    if ((rc = set_limits(steps, per_foam))) return rc;
    for (int64_t guard = 0; guard < (int64_t)1 << 40; ++guard) {
        if ((rc = read_dyn_all(d))) return rc;
        const int64_t left = steps_left(d);
        if (left <= 0) break;
        // rounds of at most kRoundGroups attempt groups, so a halted foam
        // (capacity growth) waits at most that long for the host
        const int64_t launches = std::min<int64_t>(std::max(1, kRoundGroups / groups),
                                                   std::max<int64_t>(1, (left + groups - 1) / groups));
        for (int64_t k = 0; k < launches; ++k) E_TRY(hipGraphLaunch(graph_, stream));
        if ((rc = settle(d))) return rc;
    }
    if ((rc = set_limits(0, nullptr))) return rc;
    return read_dyn_all(d);
}

int Engine::search_pending() {
    launch_search(grid());
    E_TRY(hipGetLastError());
    std::vector<Dyn> d;
    return settle(d);
}

}  // namespace cubble_amd
