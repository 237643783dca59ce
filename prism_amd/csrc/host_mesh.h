// host_mesh.h -- host-side mesh preparation for the device: geometry, face permutation,
// CSR pattern, least-squares operators, Green-Gauss sweep levels, synthetic meshes.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/prism_hip.h"

namespace prism_hip {

struct BCEntry {
    std::string kind;
    int vk{0};
    double v[3]{0, 0, 0};
};

struct HPatch {
    std::string name;
    std::map<std::string, BCEntry> bcs;  // alphabetical: the first entry decides "empty"
    bool empty{false};
    std::vector<int32_t> faces;
};

struct HFieldInfo {
    std::string name, type, grad;
    bool has_grad{false};
};

// Mesh in the permuted face order [non-empty boundary | empty boundary | interior].
struct HMesh {
    int64_t n{0}, F{0}, nb{0}, nne{0};
    std::vector<double> cx, cy, cz, vol;
    std::vector<double> sx, sy, sz, fx, fy, fz, nx, ny, nz, dx, dy, dz, gc, kx, ky, kz;
    std::vector<int32_t> owner, nbr, fpatch;
    std::vector<int32_t> cf_off, cf;
    std::vector<int32_t> rp, col, diag, slot_on, slot_no;
    // Cells that share more than one face: one CSR entry per neighbouring pair; the pair's
    // first face (in face order) writes the entry's slot, each further face an extra slot
    // rp[n] + k of the coefficient storage, and group g adds the extras
    // [dup_ptr[g], dup_ptr[g+1]) into slot dup_slot[g] in face order (linear.cpp:17-29:
    // duplicate triplets are summed).  n_extra: extra slots (0 on ordinary meshes).
    std::vector<int32_t> dup_slot, dup_ptr;
    int64_t n_extra{0};
    std::vector<int32_t> sell_off, sell_col, csr2sell;  // sliced-ELL view of the pattern
    // 16-bit column offsets: column = row + sell_base[off[s] / 64 + k] + sell_rel[slot] for
    // slot k of slice s (sellRelative); both empty when some slot's offsets span more than 16 bits
    std::vector<int16_t> sell_rel;
    std::vector<int32_t> sell_base;
    std::vector<int32_t> csc_ptr, csc_row, csc_pos;  // transpose of the pattern
    std::vector<int32_t> dia_off, csr2dia;  // DIA view (empty: the pattern does not qualify)
    std::vector<double> vtx;                 // vertex coordinates [nv][3]
    std::vector<int64_t> cell_vtx_off, cell_vtx;  // cells' own vertex lists (may be empty)
    std::vector<int32_t> v2c_off, v2c;       // vertex -> adjacent cells, ascending
    std::vector<double> pinv;  // 9 x n SoA
    std::vector<int32_t> gg_off, gg_cells;
    std::vector<int64_t> old_to_new;
    std::vector<HPatch> patches;
    std::vector<HFieldInfo> infos;
};

// Sliced-ELL view of a CSR pattern: slices of 64 rows, rows padded to the slice width
// (padding column = own row, or 0 past the last row), slot k of a row at sellPos (slots
// paired per lane).
void buildSell(int64_t n, const int32_t* rp, const int32_t* col, std::vector<int32_t>& off,
               std::vector<int32_t>& scol, std::vector<int32_t>& csr2sell);

// 16-bit column offsets of a sliced-ELL pattern: for slot k of slice s (the k-th entry of its
// 64 rows) a base b = base[off[s] / 64 + k], and for each row col - row - b in 16 bits; b is
// chosen from the span of col - row over the slot's rows, so a slot fits when that span is
// below 2^16 (on a tetrahedral box mesh the +-z neighbours, 6 nx ny rows away, share slots
// only with near neighbours).  False (rel and base empty) when a slot does not fit.
bool sellRelative(int64_t n, const std::vector<int32_t>& off, const std::vector<int32_t>& scol,
                  std::vector<int16_t>& rel, std::vector<int32_t>& base);

// Column-major view of a CSR pattern (the pattern of A^T): for column j, entries
// csc_ptr[j] .. csc_ptr[j+1]) give the row (csc_row) and the CSR position (csc_pos) of each
// A_ij, rows ascending.
void buildCsc(int64_t n, const int32_t* rp, const int32_t* col, std::vector<int32_t>& ptr,
              std::vector<int32_t>& row, std::vector<int32_t>& pos);

// Diagonal (DIA) view of a CSR pattern whose entries take at most `dmax` distinct offsets
// col - row and fill at least 3/4 of the dmax x n slots: `off` the offsets ascending,
// csr2dia[k] = d n + row for entry k on diagonal d.  False (outputs empty) otherwise.
bool buildDia(int64_t n, const int32_t* rp, const int32_t* col, int dmax, std::vector<int32_t>& off,
              std::vector<int32_t>& csr2dia);

// Builds the device-ready mesh from a descriptor (throws prism_hip::Error).
void buildMesh(const prism_mesh_desc& d, HMesh& m);
// Vertex coordinates, cell vertex lists and the vertex -> cell adjacency (part of buildMesh).
void buildVertexCells(const prism_mesh_desc& d, HMesh& m);
// Normalised inverse-distance weights of the cell -> point interpolation, one per entry of
// the vertex -> cell adjacency; fails (PRISM_ERR_MESH) on a vertex with no adjacent cell.
void idwWeights(const HMesh& m, std::vector<double>& w);

// Synthetic mesh in descriptor form (owning arrays).
struct MeshGen {
    std::vector<double> vertices;
    std::vector<int64_t> face_vtx_off, face_vtx, face_owner, face_neighbor, cell_face_off, cell_faces;
    std::vector<int64_t> cell_vtx_off, cell_vtx;
    std::vector<int32_t> face_patch;
    std::vector<std::string> patch_names;
    std::vector<const char*> patch_name_ptrs;
    int64_t n_cells{0};
    void describe(prism_mesh_desc& d);
};
void generateBox(const prism_box_params& p, MeshGen& g);

// Layout of the multicolour symmetric Gauss-Seidel preconditioner for an n-row CSR pattern
// (sorted or unsorted columns; every row stores its diagonal, diag[i] its CSR index).
// Colours: greedy in row order over the symmetrised pattern (the smallest colour no earlier
// neighbour has), at most max_colours.  Rows are renumbered colour by colour, in row order
// inside a colour: position q holds row perm[q], colour c the positions [cptr[c], cptr[c+1]).
// Off-diagonal entries as sliced ELL over positions: slices of 64 positions, slice s at
// [soff[s], soff[s+1]), entry m of position 64 s + l at soff[s] + l + 64 m, with its column in
// position numbering (scol) and its CSR index (emap); padding has emap -1 and scol = the
// position itself (0 past the last position).  dmap[q]: CSR index of position q's diagonal.
struct SgsLayout {
    int nc{0};
    std::vector<int32_t> cptr, perm, soff, scol, emap, dmap;
};
void buildSgsLayout(int64_t n, const int32_t* rp, const int32_t* col, const int32_t* diag, int max_colours,
                    SgsLayout& out);
void readUnv(const std::string& path, MeshGen& g);

}  // namespace prism_hip
