// Runs the NetCDF-3 reader (mantaray_amd/csrc/mr_netcdf3.cpp) over the files
// named on the command line, for tests/test_netcdf3_fuzz.py, which builds it
// with the address and undefined-behaviour sanitizers: each file is opened
// and every variable a bathymetry or current open reads is read as f32 and
// as f64.  Prints one status line per file and exits 0 unless a sanitizer
// or an exception stops it.
#include <cstdio>
#include <string>
#include <vector>

#include "mr_netcdf3.h"

int main(int argc, char **argv)
{
    for (int i = 1; i < argc; ++i) {
        mr::NcFile f;
        std::string msg;
        const int rc = mr::nc3_open(argv[i], f, msg);
        int reads = 0;
        if (rc == 0) {
            for (const char *name : {"x", "y", "depth", "u", "v"}) {
                std::vector<float> a;
                std::vector<double> b;
                reads += mr::nc3_read_f32(f, name, a, msg) == 0;
                reads += mr::nc3_read_f64(f, name, b, msg) == 0;
            }
        }
        std::printf("%s %d %d\n", argv[i], rc, reads);
    }
    return 0;
}
