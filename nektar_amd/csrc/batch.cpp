#include "batch.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>

namespace nektar_amd {

namespace {

double entry_work(const MFOperator &o)
{
    // dominant cost ~ points x (points per line + modes) per element
    return (double)o.nelmt * o.nqtot * (o.nq[0] + o.nm);
}

template <typename T>
void upload_items(DeviceBuffer &buf, const std::vector<GroupItem<T>> &items)
{
    buf.allocate(items.size() * sizeof(GroupItem<T>));
    NMF_HIP(hipMemcpy(buf.get(), items.data(), items.size() * sizeof(GroupItem<T>), hipMemcpyHostToDevice));
}

}  // namespace

Batch::Batch() = default;

Batch::~Batch()
{
    (void)hipDeviceSynchronize();
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (auto s : aux_) (void)hipStreamDestroy(s);
    for (auto e : join_) (void)hipEventDestroy(e);
    if (fork_) (void)hipEventDestroy(fork_);
    if (cap_) (void)hipStreamDestroy(cap_);
}

void Batch::add(MFOperator *op, const void *in, void *out0, void *out1, void *out2)
{
    NMF_REQUIRE(op, NMF_ERR_INVALID, "null operator");
    NMF_REQUIRE(op->nelmt == 0 || (in && out0), NMF_ERR_INVALID, "null input/output pointer");
    if (op->op == OpType::PhysDeriv)
        NMF_REQUIRE(op->nelmt == 0 || (out1 && out2), NMF_ERR_INVALID, "PhysDeriv needs three outputs");
    NMF_REQUIRE(op->generic || hex_supported(op->nm, op->nq[0]) || op->shape != Shape::Hex, NMF_ERR_UNSUPPORTED,
                "batches take hex operators with Q = NUMMODES + 1 or + 2");
    entries_.push_back(Entry{op, in, out0, out1, out2, op->version});
    built_ = false;
}

void Batch::set_mode(int mode)
{
    NMF_REQUIRE(mode == Graph || mode == Grouped || mode == PerItem, NMF_ERR_INVALID, "unknown batch mode");
    mode_ = mode;
}

void Batch::set_streams(int n)
{
    NMF_REQUIRE(n >= 1 && n <= 8, NMF_ERR_INVALID, "streams must be 1..8");
    if (n != nstreams_) {
        nstreams_ = n;
        built_ = false;
    }
}

bool Batch::stale() const
{
    if (!built_) return true;
    for (const auto &e : entries_)
        if (e.op->version != e.version) return true;
    return false;
}

int64_t Batch::dofs() const
{
    int64_t d = 0;
    for (const auto &e : entries_) {
        const MFOperator &o = *e.op;
        const bool coeffs = o.op == OpType::BwdTrans || o.op == OpType::Helmholtz || o.op == OpType::AdvDiffReact;
        d += (int64_t)o.nelmt * (coeffs ? o.ncoeff : o.nqtot);
    }
    return d;
}

int Batch::groups()
{
    if (stale()) build();
    return (int)groups_.size();
}

void Batch::release_graph()
{
    if (graph_) {
        NMF_HIP(hipGraphExecDestroy(graph_));
        graph_ = nullptr;
    }
}

void Batch::ensure_streams()
{
    while ((int)aux_.size() < nstreams_ - 1) {
        hipStream_t s;
        NMF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        aux_.push_back(s);
        hipEvent_t e;
        NMF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        join_.push_back(e);
    }
    if (!fork_) NMF_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    if (!cap_) NMF_HIP(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
}

void Batch::build()
{
    // item tables may be in use by earlier launches
    NMF_HIP(hipDeviceSynchronize());
    release_graph();
    groups_.clear();
    singles_.clear();
    single_stream_.clear();
    std::map<std::tuple<int, int, int, int, int, int>, int> index;
    for (int i = 0; i < (int)entries_.size(); ++i) {
        Entry &e = entries_[i];
        MFOperator &o = *e.op;
        o.prepare();
        e.version = o.version;
        if (o.nelmt == 0) continue;
        if (o.op != OpType::Helmholtz || o.generic) {  // runtime-size operators run alone
            singles_.push_back(i);
            continue;
        }
        const auto key = std::make_tuple((int)o.shape, o.nm, o.nq[0], o.tile, (int)o.fp32, (int)o.deformed);
        auto it = index.find(key);
        if (it == index.end()) {
            Group g;
            g.shape = o.shape;
            g.nm = o.nm;
            g.nq = o.nq[0];
            g.tile = o.tile;
            g.fp32 = o.fp32;
            g.deformed = o.deformed;
            groups_.push_back(std::move(g));
            it = index.emplace(key, (int)groups_.size() - 1).first;
        }
        groups_[it->second].members.push_back(i);
    }
    for (auto &g : groups_) {
        const int epb = g.shape == Shape::Hex ? hex_tile_elems(g.nq, g.tile) : tet_tile_elems(g.nq, g.tile);
        int block0 = 0;
        std::vector<GroupItem<double>> i64;
        std::vector<GroupItem<float>> i32;
        for (int m : g.members) {
            const Entry &e = entries_[m];
            const MFOperator &o = *e.op;
            const int nb = (o.nelmt + epb - 1) / epb;
            if (g.fp32)
                i32.push_back(GroupItem<float>{static_cast<const float *>(o.geo.get()),
                                               static_cast<const float *>(e.in), static_cast<float *>(e.out0),
                                               (float)o.lambda, o.nelmt, block0});
            else
                i64.push_back(GroupItem<double>{static_cast<const double *>(o.geo.get()),
                                                static_cast<const double *>(e.in), static_cast<double *>(e.out0),
                                                o.lambda, o.nelmt, block0});
            NMF_REQUIRE((int64_t)block0 + nb < (1ll << 31), NMF_ERR_INVALID, "batch kind exceeds 2^31 workgroups");
            block0 += nb;
            g.work += entry_work(o);
        }
        g.nblocks = block0;
        if (g.fp32)
            upload_items(g.items, i32);
        else
            upload_items(g.items, i64);
    }
    // largest first onto the least-loaded stream
    struct Unit {
        double work;
        int group, single;
    };
    std::vector<Unit> units;
    for (int g = 0; g < (int)groups_.size(); ++g) units.push_back({groups_[g].work, g, -1});
    for (int s = 0; s < (int)singles_.size(); ++s) units.push_back({entry_work(*entries_[singles_[s]].op), -1, s});
    std::stable_sort(units.begin(), units.end(), [](const Unit &a, const Unit &b) { return a.work > b.work; });
    std::vector<double> load(nstreams_, 0.0);
    single_stream_.assign(singles_.size(), 0);
    for (const auto &u : units) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += u.work;
        if (u.group >= 0)
            groups_[u.group].stream = k;
        else
            single_stream_[u.single] = k;
    }
    built_ = true;
}

void Batch::launch_all(hipStream_t s)
{
    const bool fork = nstreams_ > 1 && groups_.size() + singles_.size() > 1;
    if (fork) {
        ensure_streams();
        NMF_HIP(hipEventRecord(fork_, s));
        for (int k = 0; k < nstreams_ - 1; ++k) NMF_HIP(hipStreamWaitEvent(aux_[k], fork_, 0));
    }
    auto stream_of = [&](int k) { return (!fork || k == 0) ? s : aux_[k - 1]; };
    for (const auto &g : groups_) {
        const MFOperator &first = *entries_[g.members[0]].op;
        GroupArgs a{g.fp32, g.deformed, (int)g.members.size(), g.items.get(), g.nblocks, first.tab.get(),
                    stream_of(g.stream), g.tile};
        hipError_t e = g.shape == Shape::Hex ? launch_hex_grouped(g.nm, g.nq, a) : launch_tet_grouped(g.nm, g.nq, a);
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("grouped launch failed: ") + hipGetErrorString(e));
    }
    for (size_t i = 0; i < singles_.size(); ++i) {
        const Entry &e = entries_[singles_[i]];
        e.op->apply(e.in, e.out0, e.out1, e.out2, stream_of(single_stream_[i]));
    }
    if (fork) {
        for (int k = 0; k < nstreams_ - 1; ++k) {
            NMF_HIP(hipEventRecord(join_[k], aux_[k]));
            NMF_HIP(hipStreamWaitEvent(s, join_[k], 0));
        }
    }
}

void Batch::run(hipStream_t s)
{
    if (stale()) build();
    if (mode_ == PerItem) {
        for (const auto &e : entries_) e.op->apply(e.in, e.out0, e.out1, e.out2, s);
        return;
    }
    if (mode_ == Grouped) {
        launch_all(s);
        return;
    }
    if (!graph_) {
        ensure_streams();
        hipGraph_t g;
        NMF_HIP(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
        try {
            launch_all(cap_);
        } catch (...) {
            hipGraph_t dead;
            (void)hipStreamEndCapture(cap_, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        NMF_HIP(hipStreamEndCapture(cap_, &g));
        hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    NMF_HIP(hipGraphLaunch(graph_, s));
}

void Batch::autotune(hipStream_t s)
{
    // one choice per kind (shape, order, points, precision, geometry),
    // timed on the kind's largest operator
    std::map<std::tuple<int, int, int, int, int>, std::vector<MFOperator *>> kinds;
    for (const auto &e : entries_) {
        MFOperator *o = e.op;
        if (o->op != OpType::Helmholtz || o->nelmt == 0 || o->tile_candidates().empty()) continue;
        auto &v = kinds[std::make_tuple((int)o->shape, o->nm, o->nq[0], (int)o->fp32, (int)o->deformed)];
        if (std::find(v.begin(), v.end(), o) == v.end()) v.push_back(o);
    }
    for (auto &kv : kinds) {
        auto &ops = kv.second;
        MFOperator *big = *std::max_element(ops.begin(), ops.end(),
                                            [](const MFOperator *a, const MFOperator *b) { return a->nelmt < b->nelmt; });
        const int t = big->autotune(s);
        for (MFOperator *o : ops)
            if (o != big) o->set_tile(t);
    }
}

}  // namespace nektar_amd
