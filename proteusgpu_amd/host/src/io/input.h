// Parameter file and initial conditions (reference src/io/input.h:11-54):
// param.txt "key = value" lines with '#' comments and typed getters that
// throw std::runtime_error; the IC file layout /header@dimension, /seedpos,
// /rho, /vel, /Energy (input.cpp:121-310).
#ifndef PGPU_HOST_INPUT_H
#define PGPU_HOST_INPUT_H

#include <map>
#include <string>
#include <vector>

#include "../global/allvars.h"

struct ICHeader {
    int dimension;
};

struct ICData {
    std::vector<double> seedpos;       // dimension * numSeeds
    std::vector<hsize_t> seedpos_dims; // [numSeeds, dimension]
    std::vector<double> rho;
    std::vector<double> vel;           // dimension * numSeeds
    std::vector<double> Energy;        // total energy per volume
    ICHeader header;
};

class InputHandler {
   public:
    explicit InputHandler(const std::string& filename = "param.txt");
    bool loadParameters();
    bool has(const std::string& key) const;
    std::string getParameter(const std::string& key) const;
    int getParameterInt(const std::string& key) const;
    double getParameterDouble(const std::string& key) const;
    bool getParameterBool(const std::string& key) const;
    // optional keys with defaults (keys the reference hard-codes)
    double getDouble(const std::string& key, double def) const;
    int getInt(const std::string& key, int def) const;
    bool getBool(const std::string& key, bool def) const;
    std::string getString(const std::string& key, const std::string& def) const;
#ifdef USE_HDF5
    bool readICFile(const std::string& filename, ICData& ic);
#endif
   private:
    std::map<std::string, std::string> parameters;
    std::string paramFilePath;
};

extern InputHandler input;
extern ICData icData;

#endif
