// C entry points over the reference's own CPU ops (test infrastructure only).
//
// oracle/ref.mk compiles torchCompactRadius 0.5.5's
// src/torchCompactRadius/cppSrc/neighborhoodDynamic_cpu.cpp and hashing_cpu.cpp
// where they lie, together with this file, into oracle/_ref/libtcr_ref.so.
// Each function below wraps the caller's plain host buffers as CPU tensors
// (no copy), calls the reference op with the argument list of
// cpp/neighborhood.h:102-124 / cpp/hashing.h:38, and copies the op's output
// tensor back.  The tests use it to check the oracle's restatement and the
// HIP library's reference-layout ops against the reference's own code; the
// product never loads it.
#include <cstring>
#include <string>

#include "hashing.h"
#include "neighborhood.h"

namespace {
thread_local std::string g_err;

torch::Tensor wrap_f32(const float *p, int64_t n0, int64_t n1 = -1) {
    auto opt = torch::TensorOptions().dtype(torch::kFloat32);
    if (n1 < 0) return torch::from_blob(const_cast<float *>(p), {n0}, opt);
    return torch::from_blob(const_cast<float *>(p), {n0, n1}, opt);
}
torch::Tensor wrap_i32(const int32_t *p, int64_t n0, int64_t n1 = -1) {
    auto opt = torch::TensorOptions().dtype(torch::kInt32);
    if (n1 < 0) return torch::from_blob(const_cast<int32_t *>(p), {n0}, opt);
    return torch::from_blob(const_cast<int32_t *>(p), {n0, n1}, opt);
}
const char *mode_name(int mode) {
    switch (mode) {
        case 0: return "symmetric";
        case 1: return "gather";
        case 2: return "scatter";
        case 3: return "superSymmetric";
        default: return "invalid";
    }
}
}  // namespace

extern "C" {

const char *ref_last_error(void) { return g_err.c_str(); }

// torch::set_num_threads for the ops' at::parallel_for (0 leaves the default).
void ref_set_threads(int n) {
    if (n > 0) at::set_num_threads(n);
}

// countNeighbors (cpp/neighborhoodDynamic_cpu.cpp:495-519).  Supports may be
// NULL (an undefined tensor) in the mode that does not read them; periodic is
// one byte per dimension (torch.bool).  Returns 0, or -1 with the op's
// exception message in ref_last_error().
int ref_count_neighbors(const float *qpos, const float *qsup, int64_t nq, int32_t dim, int32_t search_range,
                        const float *spos, const float *ssup, int64_t nr, const int32_t *hash_table, int64_t L,
                        const int32_t *num_cells, const int32_t *cell_table, int64_t ncells, double hcell,
                        const float *dmin, const float *dmax, const uint8_t *periodic, int32_t mode,
                        int32_t *counts) {
    try {
        auto per = torch::from_blob(const_cast<uint8_t *>(periodic), {dim}, torch::TensorOptions().dtype(torch::kBool));
        auto c = TORCH_EXTENSION_NAME::countNeighbors(
            wrap_f32(qpos, nq, dim), qsup ? wrap_f32(qsup, nq) : torch::Tensor(), search_range,
            wrap_f32(spos, nr, dim), ssup ? wrap_f32(ssup, nr) : torch::Tensor(), wrap_i32(hash_table, L, 2), L,
            wrap_i32(num_cells, dim), wrap_i32(cell_table, ncells, 3), wrap_f32(dmin, dim), hcell,
            wrap_f32(dmax, dim), wrap_f32(dmin, dim), per, mode_name(mode), false);
        std::memcpy(counts, c.data_ptr<int32_t>(), sizeof(int32_t) * nq);
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// buildNeighborList (cpp/neighborhoodDynamic_cpu.cpp:521-547): int32 counts
// and offsets, list_length entries of (i, j_sorted) int64.
int ref_build_neighbor_list(const int32_t *counts, const int32_t *offsets, int64_t list_length, const float *qpos,
                            const float *qsup, int64_t nq, int32_t dim, int32_t search_range, const float *spos,
                            const float *ssup, int64_t nr, const int32_t *hash_table, int64_t L,
                            const int32_t *num_cells, const int32_t *cell_table, int64_t ncells, double hcell,
                            const float *dmin, const float *dmax, const uint8_t *periodic, int32_t mode,
                            int64_t *list_i, int64_t *list_j) {
    try {
        auto per = torch::from_blob(const_cast<uint8_t *>(periodic), {dim}, torch::TensorOptions().dtype(torch::kBool));
        auto r = TORCH_EXTENSION_NAME::buildNeighborList(
            wrap_i32(counts, nq), wrap_i32(offsets, nq), list_length, wrap_f32(qpos, nq, dim),
            qsup ? wrap_f32(qsup, nq) : torch::Tensor(), search_range, wrap_f32(spos, nr, dim),
            ssup ? wrap_f32(ssup, nr) : torch::Tensor(), wrap_i32(hash_table, L, 2), L, wrap_i32(num_cells, dim),
            wrap_i32(cell_table, ncells, 3), wrap_f32(dmin, dim), hcell, wrap_f32(dmax, dim), wrap_f32(dmin, dim),
            per, mode_name(mode), false);
        std::memcpy(list_i, std::get<0>(r).data_ptr<int64_t>(), sizeof(int64_t) * list_length);
        std::memcpy(list_j, std::get<1>(r).data_ptr<int64_t>(), sizeof(int64_t) * list_length);
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// computeHashIndices (cpp/hashing_cpu.cpp:52-56): cells int32[n, dim] -> int32[n].
int ref_compute_hash_indices(const int32_t *cells, int64_t n, int32_t dim, int64_t L, int32_t *hash) {
    try {
        auto h = TORCH_EXTENSION_NAME::computeHashIndices(wrap_i32(cells, n, dim), L);
        std::memcpy(hash, h.data_ptr<int32_t>(), sizeof(int32_t) * n);
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
