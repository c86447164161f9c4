#include "input.h"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <stdexcept>

InputHandler input;
ICData icData;

namespace {
std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
std::runtime_error missing(const std::string& key) {
    return std::runtime_error("Error: Required parameter '" + key + "' not found in parameter file");
}
}  // namespace

InputHandler::InputHandler(const std::string& filename) : paramFilePath(filename) {}

// input.cpp:21-64
bool InputHandler::loadParameters() {
    std::ifstream f(paramFilePath);
    if (!f.is_open()) {
        std::cerr << "INPUT: Error! Could not open parameter file: " << paramFilePath << std::endl;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        size_t c = val.find('#');
        if (c != std::string::npos) val = trim(val.substr(0, c));
        parameters[key] = val;
    }
    std::cout << "INPUT: Loaded " << parameters.size() << " parameters from " << paramFilePath << std::endl;
    return true;
}

bool InputHandler::has(const std::string& key) const { return parameters.count(key) != 0; }

std::string InputHandler::getParameter(const std::string& key) const {
    auto it = parameters.find(key);
    if (it == parameters.end()) throw missing(key);
    return it->second;
}

int InputHandler::getParameterInt(const std::string& key) const {
    std::string v = getParameter(key);
    try {
        return std::stoi(v);
    } catch (const std::exception&) {
        throw std::runtime_error("Error: Could not convert parameter '" + key + "' with value '" + v + "' to int");
    }
}

double InputHandler::getParameterDouble(const std::string& key) const {
    std::string v = getParameter(key);
    try {
        return std::stod(v);
    } catch (const std::exception&) {
        throw std::runtime_error("Error: Could not convert parameter '" + key + "' with value '" + v +
                                 "' to double");
    }
}

bool InputHandler::getParameterBool(const std::string& key) const {
    std::string v = getParameter(key), l = v;
    std::transform(l.begin(), l.end(), l.begin(), ::tolower);
    if (l == "true" || l == "1" || l == "yes" || l == "on") return true;
    if (l == "false" || l == "0" || l == "no" || l == "off") return false;
    throw std::runtime_error("Error: Could not convert parameter '" + key + "' with value '" + v +
                             "' to bool (expected: true/false/1/0/yes/no/on/off)");
}

double InputHandler::getDouble(const std::string& k, double d) const { return has(k) ? getParameterDouble(k) : d; }
int InputHandler::getInt(const std::string& k, int d) const { return has(k) ? getParameterInt(k) : d; }
bool InputHandler::getBool(const std::string& k, bool d) const { return has(k) ? getParameterBool(k) : d; }
std::string InputHandler::getString(const std::string& k, const std::string& d) const {
    return has(k) ? getParameter(k) : d;
}

#ifdef USE_HDF5
namespace {
bool read_dataset(hid_t file, const char* name, std::vector<double>& out, std::vector<hsize_t>* dims) {
    hid_t ds = H5Dopen(file, name, H5P_DEFAULT);
    if (ds < 0) {
        std::cerr << "INPUT: Error! Could not open " << name << " dataset" << std::endl;
        return false;
    }
    hid_t sp = H5Dget_space(ds);
    int nd = H5Sget_simple_extent_ndims(sp);
    std::vector<hsize_t> d(nd);
    H5Sget_simple_extent_dims(sp, d.data(), nullptr);
    hsize_t total = 1;
    for (hsize_t x : d) total *= x;
    out.resize(total);
    herr_t st = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data());
    H5Sclose(sp);
    H5Dclose(ds);
    if (dims) *dims = d;
    return st >= 0;
}
}  // namespace

// input.cpp:121-310: /header@dimension, /seedpos[N,d], /rho[N], /vel[N,d], /Energy[N]
bool InputHandler::readICFile(const std::string& filename, ICData& ic) {
    hid_t file = H5Fopen(filename.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) {
        std::cerr << "INPUT: Error! Could not open IC file: " << filename << std::endl;
        return false;
    }
    hid_t hdr = H5Gopen(file, "header", H5P_DEFAULT);
    hid_t at = hdr >= 0 ? H5Aopen(hdr, "dimension", H5P_DEFAULT) : -1;
    bool ok = at >= 0 && H5Aread(at, H5T_NATIVE_INT, &ic.header.dimension) >= 0;
    if (at >= 0) H5Aclose(at);
    if (hdr >= 0) H5Gclose(hdr);
    if (!ok) {
        std::cerr << "INPUT: Error! Could not read dimension attribute from IC file" << std::endl;
        H5Fclose(file);
        return false;
    }
    if (ic.header.dimension != DIMENSION) {
        std::cerr << "INPUT: Error! IC file dimension mismatch!" << std::endl;
        H5Fclose(file);
        return false;
    }
    ok = read_dataset(file, "seedpos", ic.seedpos, &ic.seedpos_dims) && read_dataset(file, "rho", ic.rho, nullptr) &&
         read_dataset(file, "vel", ic.vel, nullptr) && read_dataset(file, "Energy", ic.Energy, nullptr);
    H5Fclose(file);
    if (ok) std::cout << "INPUT: Read " << ic.seedpos_dims[0] << " seeds from " << filename << std::endl;
    return ok;
}
#endif
