// voronoi:: interface of the reference (src/voronoi/voronoi.h:22-81,
// periodic_mesh.h:22) on libpgpu.  The VMesh is the reference's struct of
// arrays in original input order; the device mesh behind it stays in HBM and
// is what hydro::hydro_step runs on.
#ifndef PGPU_HOST_VORONOI_H
#define PGPU_HOST_VORONOI_H

#include <vector>

#include "../global/allvars.h"
#include "../io/meshdata.h"
#include "../knn/knn.h"

struct VMesh {
    hsize_t* cell_ids;      // cell ids
    double3* seeds;         // seed points (z = 0 in 2D)
    hsize_t n_seeds;        // number of cells
    double* volumes;        // area in 2D, volume in 3D
    hsize_t* face_counts;   // faces per cell
    hsize_t* face_ptr;      // start of each cell's faces
    int* neighbor_cell;     // original id of the neighbour, -1 = box plane
    double3* face_normal;   // unit (seed_i - seed_j)
    double* face_area;      // edge length in 2D, area in 3D
    hsize_t num_faces;
    hsize_t n_hydro;        // active cells (n_seeds - n_hydro ghosts)
    hsize_t* ghost_ids;     // original cell of each ghost
    // no reference counterpart: per-cell status and minimal certified k
    int* status;
    int* kcert;
    pg_mesh* dev;           // the device mesh
};

namespace voronoi {

// voronoi.h:64: cell arrays for n_seeds and face arrays of the given
// capacity (compute_cells resizes them to the mesh it computes)
VMesh* allocate_vmesh(hsize_t n_seeds, hsize_t initial_face_capacity);
void free_vmesh(VMesh* mesh);                                          // voronoi.h:65
VMesh* compute_mesh(POINT_TYPE* pts_data, int num_points);             // voronoi.h:68
VMesh* compute_periodic_mesh(POINT_TYPE* pts_data, hsize_t num_points);  // periodic_mesh.h:22
// no reference counterpart: the unit box as reflective walls
VMesh* compute_reflective_mesh(POINT_TYPE* pts_data, hsize_t num_points);
// voronoi.h:69: the cells of a solved KNN problem (no restarts), in its
// sorted order with sorted neighbour ids; stat per sorted seed
void compute_cells(int N_seedpts, knn_problem* knn, std::vector<Status>& stat, VMesh* mesh);
// voronoi.h:77: cells of compute_cells back to input order (sorted_to_original
// must be the KNN problem's permutation; the reorder runs on the device)
void unpermute_vmesh(VMesh* mesh, const unsigned int* sorted_to_original);
// voronoi.h:81, voronoi.cpp:336-421: the snapshot form of a mesh (header
// n = n_hydro cells, k = _K_, nmax = _MAX_P_, extent 1, seed 0; cell arrays
// of the n_hydro cells, face arrays of every cell)
void vmesh_to_meshdata(VMesh* mesh, MeshCellData& meshData);

}  // namespace voronoi

#endif
