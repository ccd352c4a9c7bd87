// mr_netcdf3.h -- minimal NetCDF-3 reader (classic CDF-1 and 64-bit-offset
// CDF-2), the input format of the reference's gridded fields.
//
// The reference reads its files through the netcdf3 crate 0.5.2
// (src/bathymetry/cartesian_netcdf3.rs:167-256,
// src/current/cartesian_current.rs:58-213): a variable is read whole, in file
// order, with no use of attributes (_FillValue, scale_factor) or of the
// dimension order.  This reader does the same and follows the published
// NetCDF classic format specification.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mr {

enum class NcType : int32_t { Byte = 1, Char = 2, Short = 3, Int = 4, Float = 5, Double = 6 };

struct NcVar {
    std::string name;
    NcType type;
    std::vector<int32_t> dimids;
    bool is_record = false;
    uint64_t begin = 0;
    uint64_t vsize = 0;
    uint64_t n_per_record = 1;     // elements per record (all elements if not a record var)
};

struct NcFile {
    int version = 0;
    uint64_t numrecs = 0;
    uint64_t recsize = 0;
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<NcVar> vars;
    std::vector<uint8_t> bytes;    // the whole file
};

// Returns 0 on success, -2 for an I/O error, -3 for a format error; msg holds
// a description on failure.
int nc3_open(const std::string &path, NcFile &f, std::string &msg);

// Reads variable `name` converted element-wise with C casts, which are the
// Rust `as` conversions for these types.  Returns 0, or -3 if the variable
// is missing.
int nc3_read_f32(const NcFile &f, const std::string &name, std::vector<float> &out,
                 std::string &msg);
int nc3_read_f64(const NcFile &f, const std::string &name, std::vector<double> &out,
                 std::string &msg);

}  // namespace mr
