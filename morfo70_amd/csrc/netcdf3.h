// netcdf3.h -- self-contained NetCDF classic-format reader and writer (host code).
//
// The reference reads its bathymetry and resume files and writes its m70_*.nc snapshots
// through netcdf-c (class MorfoIO::Netcdf, include/Netcdf.h, src/Netcdf.cpp), which this
// environment does not have. The files involved use only fixed-size dimensions, global
// attributes and non-record numeric variables, so this implements that part of the
// classic format (CDF-1, and CDF-2 with 64-bit offsets for files past 2 GiB): big-endian
// header (dimensions, global attributes, variables with their attributes, type, size and
// offset) followed by the variables' data, each padded to four bytes.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace m70nc {

enum Type { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

struct Attr {
    std::string name;
    int type = NC_DOUBLE;
    std::vector<double> values;   // numeric attributes
    std::string text;             // NC_CHAR attributes
};

struct Var {
    std::string name;
    std::vector<int> dims;        // dimension ids
    std::vector<Attr> atts;
    int type = NC_DOUBLE;
    uint64_t vsize = 0;
    uint64_t begin = 0;
};

struct File {
    int version = 1;
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<Attr> gatts;
    std::vector<Var> vars;
    std::string path;
};

// Parses the header; false with a message on failure.
bool read_header(const std::string &path, File &f, std::string &err);
// Reads variable `name` converted to double (n = its element count); false if absent.
bool read_var(const File &f, const std::string &name, std::vector<double> &out, std::string &err);
// Global attribute `name`, first value as double.
bool read_att(const File &f, const std::string &name, double &out, std::string &err);

// A variable to write: double data over dimension ids.
struct OutVar {
    std::string name;
    std::vector<int> dims;
    const double *data;
};
// Writes dimensions, double global attributes and double variables; CDF-1 when every
// offset fits in 31 bits, CDF-2 otherwise (or when force_64bit).
bool write_file(const std::string &path, const std::vector<std::pair<std::string, uint64_t>> &dims,
                const std::vector<std::pair<std::string, double>> &gatts, const std::vector<OutVar> &vars,
                bool force_64bit, std::string &err);

}  // namespace m70nc
