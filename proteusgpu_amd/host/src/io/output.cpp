#include "output.h"

#include <sys/stat.h>

#include <iostream>
#include <vector>

OutputHandler output;

bool OutputHandler::initialize() {
    struct stat st;
    if (stat(outputDirectory.c_str(), &st) == 0) return true;
    if (mkdir(outputDirectory.c_str(), 0755) != 0) {
        std::cerr << "OUTPUT: Error! Could not create output directory: " << outputDirectory << std::endl;
        return false;
    }
    return true;
}

#ifdef USE_HDF5
namespace {
// One dataset of n0 (x n1) elements; like the reference's writer
// (output.cpp:95-226, 239-290), an empty or missing array writes no dataset.
template <class T>
void write(hid_t group, const char* name, hid_t type, const T* data, hsize_t n0, hsize_t n1 = 0) {
    if (n0 == 0 || data == nullptr) return;
    hsize_t dims[2] = {n0, n1};
    hid_t sp = H5Screate_simple(n1 ? 2 : 1, dims, nullptr);
    hid_t ds = H5Dcreate(group, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (n0) H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    H5Sclose(sp);
}
template <class T>
void attr(hid_t g, const char* name, hid_t type, T v) {
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate(g, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, type, &v);
    H5Aclose(a);
    H5Sclose(sp);
}
}  // namespace

bool OutputHandler::writeMeshFile(const std::string& filename, const MeshCellData& md, const primvars* p,
                                  int n_hydro) {
    std::string path = outputDirectory + filename;  // output.cpp:29 joins without a separator
    hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) {
        std::cerr << "OUTPUT: Error! Could not create file: " << path << std::endl;
        return false;
    }
    const int n = md.header.n;
    hid_t h = H5Gcreate(f, "header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    attr(h, "dimension", H5T_NATIVE_INT, md.header.dimension);
    attr(h, "extent", H5T_NATIVE_DOUBLE, md.header.extent);
    attr(h, "n", H5T_NATIVE_INT, n);
    attr(h, "k", H5T_NATIVE_INT, md.header.k);
    attr(h, "nmax", H5T_NATIVE_INT, md.header.nmax);
    attr(h, "seed", H5T_NATIVE_INT, md.header.seed);
    hbool_t store = md.header.store_edge_coords ? 1 : 0;
    attr(h, "store_edge_coords", H5T_NATIVE_HBOOL, store);
    H5Gclose(h);
    hid_t cells = H5Gcreate(f, "cells", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    write(cells, "cell_ids", H5T_NATIVE_INT, md.cell_ids.data(), n);
    write(cells, "seeds", H5T_NATIVE_DOUBLE, md.seeds.data(), n, DIMENSION);
    write(cells, "volumes", H5T_NATIVE_DOUBLE, md.volumes.data(), n);
    write(cells, "face_counts", H5T_NATIVE_INT, md.face_counts.data(), n);
    hid_t faces = H5Gcreate(cells, "faces", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const hsize_t nf = md.faces.area.size();
    write(faces, "neighbor_cell", H5T_NATIVE_INT, md.faces.neighbor_cell.data(), nf);
    write(faces, "normal", H5T_NATIVE_DOUBLE, md.faces.normal.data(), nf, DIMENSION);
    write(faces, "area", H5T_NATIVE_DOUBLE, md.faces.area.data(), nf);
    H5Gclose(faces);
    H5Gclose(cells);
    // output.cpp:236-293: the hydro group always, each array only when present
    hid_t hy = H5Gcreate(f, "hydro", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (p) {
        write(hy, "rho", H5T_NATIVE_DOUBLE, p->rho, n_hydro);
        write(hy, "vel", H5T_NATIVE_DOUBLE, (const double*)p->v, n_hydro, DIMENSION);
        write(hy, "Energy", H5T_NATIVE_DOUBLE, p->E, n_hydro);
    }
    H5Gclose(hy);
    H5Fclose(f);
    std::cout << "OUTPUT: wrote " << path << std::endl;
    return true;
}

bool OutputHandler::writeMeshFile(const std::string& filename, const VMesh* m, const primvars* p, int n_hydro) {
    MeshCellData md;
    voronoi::vmesh_to_meshdata(const_cast<VMesh*>(m), md);
    return writeMeshFile(filename, md, p, n_hydro);
}

bool OutputHandler::writeKNNFile(const std::string& filename, const double* points, const unsigned int* nearest,
                                 const unsigned int* permutation, int num_points, int k) {
    std::string path = outputDirectory + filename;  // output.cpp:312 joins without a separator
    hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) {
        std::cerr << "OUTPUT: Error! Could not create file: " << path << std::endl;
        return false;
    }
    hid_t h = H5Gcreate(f, "header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    attr(h, "dimension", H5T_NATIVE_INT, (int)DIMENSION);
    attr(h, "num_points", H5T_NATIVE_INT, num_points);
    attr(h, "k", H5T_NATIVE_INT, k);
    H5Gclose(h);
    hid_t g = H5Gcreate(f, "knn", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    write(g, "points", H5T_NATIVE_DOUBLE, points, (hsize_t)num_points, DIMENSION);
    write(g, "nearest", H5T_NATIVE_UINT, nearest, (hsize_t)num_points, (hsize_t)k);
    write(g, "permutation", H5T_NATIVE_UINT, permutation, (hsize_t)num_points);
    H5Gclose(g);
    H5Fclose(f);
    std::cout << "OUTPUT: wrote " << path << std::endl;
    return true;
}
#endif
