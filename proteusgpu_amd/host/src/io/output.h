// Snapshot writer with the reference's HDF5 layout (src/io/output.cpp:28-303):
// /header@{dimension,extent,n,k,nmax,seed,store_edge_coords},
// /cells/{cell_ids,seeds,volumes,face_counts}, /cells/faces/{neighbor_cell,normal,area},
// /hydro/{rho,vel,Energy}.  Cell arrays cover n_hydro cells while the face
// arrays cover every cell, ghosts included (voronoi.cpp:341,387).
#ifndef PGPU_HOST_OUTPUT_H
#define PGPU_HOST_OUTPUT_H

#include <string>

#include "../global/allvars.h"
#include "../voronoi/voronoi.h"
#include "meshdata.h"

class OutputHandler {
   public:
    explicit OutputHandler(const std::string& dir = "./output/") : outputDirectory(dir) {}
    bool initialize();
    std::string getOutputDirectory() const { return outputDirectory; }
#ifdef USE_HDF5
    // output.h:74: the snapshot of vmesh_to_meshdata's result
    bool writeMeshFile(const std::string& filename, const MeshCellData& meshData, const primvars* primvar,
                       int n_hydro);
    // the same file straight from a VMesh (vmesh_to_meshdata + the above)
    bool writeMeshFile(const std::string& filename, const VMesh* mesh, const primvars* primvar, int n_hydro);
    // KNN dump (output.cpp:305-411): /header@{dimension,num_points,k},
    // /knn/{points[num_points][d], nearest[num_points][k], permutation[num_points]},
    // the first num_points rows of the sorted problem.
    bool writeKNNFile(const std::string& filename, const double* points, const unsigned int* nearest,
                      const unsigned int* permutation, int num_points, int k);
#endif
   private:
    std::string outputDirectory;
};

extern OutputHandler output;

#endif
