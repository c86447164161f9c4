// Snapshot data of one mesh, with the reference's struct and field names
// (src/io/output.h:13-48): what vmesh_to_meshdata fills and writeMeshFile
// writes.  Cell arrays cover the n hydro cells; face arrays every cell.
#ifndef PGPU_HOST_MESHDATA_H
#define PGPU_HOST_MESHDATA_H

#include <vector>

#include "../global/allvars.h"

struct MeshHeader {
    int dimension = DIMENSION;
    double extent = 1.0;
    int n = 0;
    int k = 0;
    int nmax = 0;
    int seed = 0;
    bool store_edge_coords = false;
};

struct MeshFaceData {
    std::vector<int> neighbor_cell;
    std::vector<double> normal;        // numFaces x dimension
    std::vector<hsize_t> normal_dims;  // [numFaces, dimension]
    std::vector<double> area;
};

struct MeshCellData {
    MeshHeader header;
    std::vector<int> cell_ids;
    std::vector<double> seeds;        // numCells x dimension
    std::vector<hsize_t> seeds_dims;  // [numCells, dimension]
    std::vector<double> volumes;
    std::vector<int> face_counts;
    MeshFaceData faces;
};

#endif
