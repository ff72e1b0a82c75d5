// Batched application of many matrix-free operators of uneven kinds and
// sizes (the Collections loop of ExpList::GeneralMatrixOp,
// MultiRegions/ExpList.cpp:2666-2757, applied to a whole batch at once).
//
// Helmholtz entries of one kind (shape, modes, points, precision, geometry
// type, tile) run as one grouped launch over a device item table; other
// entries launch on their own.  Kinds are spread over a few streams,
// largest first onto the least-loaded stream, and the whole batch is
// captured once into a hipGraph that later runs replay.  Any change to an
// entry's operator (geometry, factors, tile) is picked up at the next run.
#pragma once

#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "mf_operator.hpp"

namespace nektar_amd {

class Batch {
public:
    enum Mode { Graph = 0, Grouped = 1, PerItem = 2 };

    Batch();
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    // The operator must outlive the batch; in/out are device arrays of the
    // operator's precision (out1, out2 only for PhysDeriv).
    void add(MFOperator *op, const void *in, void *out0, void *out1, void *out2);
    void set_mode(int mode);
    void set_streams(int n);
    void run(hipStream_t s);
    // Autotune each Helmholtz hex kind on its largest member and give the
    // choice to every member of the kind.
    void autotune(hipStream_t s);

    int items() const { return (int)entries_.size(); }
    int groups();
    int64_t dofs() const;  // sum of element-local input DOFs

private:
    struct Entry {
        MFOperator *op;
        const void *in;
        void *out0, *out1, *out2;
        uint64_t version;
    };
    struct Group {
        Shape shape;
        int nm, nq, tile;
        bool fp32, deformed;
        std::vector<int> members;
        DeviceBuffer items;
        int nblocks = 0;
        double work = 0.0;
        int stream = 0;
    };

    bool stale() const;
    void build();
    void launch_all(hipStream_t s);
    void release_graph();
    void ensure_streams();

    std::vector<Entry> entries_;
    std::vector<Group> groups_;
    std::vector<int> singles_;       // entries that launch on their own
    std::vector<int> single_stream_;
    int mode_ = Graph;
    int nstreams_ = 4;
    bool built_ = false;
    hipGraphExec_t graph_ = nullptr;
    std::vector<hipStream_t> aux_;
    std::vector<hipEvent_t> join_;
    hipEvent_t fork_ = nullptr;
    hipStream_t cap_ = nullptr;  // capture stream (the caller's may be the null stream)
};

}  // namespace nektar_amd
