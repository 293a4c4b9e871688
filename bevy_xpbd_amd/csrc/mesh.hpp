// Host-side topology for the XPBD cloth path: ClothMeshData and its builders
// (reference: src/mesh_prep.rs).  Plain C++; no device code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xpbd.h"

namespace xpbd {

struct Constraint {
    uint32_t i, j;
    float rest, comp;
};

// ClothMeshData (mesh_prep.rs:103-131) plus the batch fields (ground flag and
// per-cloth vertex offsets) of the multi-cloth layout.
struct Mesh {
    std::vector<float> positions, rest_positions, normals, uvs;
    std::vector<uint32_t> indices, tri_edge_neighbors;
    std::vector<float> inv_mass;
    std::vector<uint32_t> neighbor_other, neighbor_constraint_id, neighbor_offsets;
    std::vector<float> neighbor_rest_len, neighbor_compliance;
    std::vector<uint32_t> constraint_i, constraint_j;
    std::vector<float> constraint_rest, constraint_comp;
    std::vector<uint32_t> batch_offsets;
    std::vector<uint32_t> ground;
    std::vector<uint32_t> cloth_offsets;
    std::vector<uint32_t> cloth_cols;   // grid columns per cloth (empty: not a grid mesh)

    size_t num_particles() const { return inv_mass.size(); }
    size_t num_constraints() const { return constraint_i.size(); }
    size_t num_tris() const { return indices.size() / 3; }
};

Mesh* grid_cloth_hanging(uint32_t quad_cols, uint32_t quad_rows, float cell_size);
Mesh* parse_welded_obj(const std::string& text);
Mesh* sheet_batch(const xpbd_sheet_desc* descs, uint32_t count);
// ClothSimConfig's topology arrays (cloth_compute.rs:142-179), validated;
// nullptr with the reason in `why` and an XPBD_E_* status in `code` when an
// array is inconsistent or the device preparation fails.
Mesh* mesh_from_arrays(const xpbd_mesh_desc& d, std::string& why, int& code);
// The colouring, colour sort and CSR of assemble_constraints computed on the
// GPU (topology.hip), same results; `rows` is consumed.  Returns XPBD_OK or an
// XPBD_E_* status with the reason in `why`.
int topology_on_device(size_t n, const std::vector<uint32_t>& cloth_offsets, std::vector<Constraint>& rows, Mesh& m,
                       std::string& why);

// Greedy edge colouring + colour-sorted order + CSR cursor scatter
// (mesh_prep.rs:153-221, 584-635).  `constraints` is consumed.
void assemble_constraints(Mesh& m, size_t n, std::vector<Constraint>&& constraints);
// The CSR of m's constraint rows by cursor scatter in row order
// (mesh_prep.rs:588-634).
void scatter_csr(Mesh& m, size_t n);

void derive_collision_grid(const float* rest_vec4, uint32_t n, float thickness, xpbd_coll_grid* g);

}  // namespace xpbd
