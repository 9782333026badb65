// Drop-in process entry: `main_<system>.ex inputs`, the counterpart of the
// reference's Source/main.cpp:11-48 (ZoomyAmr solver; InitData(); Evolve();).
// The deck is the file named on the command line (run_case.py:419-422);
// plotfiles go to the working directory, the Initial/Step/Total lines to
// stdout.  Any error prints the message and exits non-zero, as amrex::Abort
// does, so run_case's check=True turns it into a Python exception.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "zamr.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s inputs\n", argv[0]);
        return 2;
    }
    std::ifstream f(argv[1]);
    if (!f) {
        std::fprintf(stderr, "cannot open inputs file %s\n", argv[1]);
        return 2;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    zamr_ctx* ctx = nullptr;
    int rc = zamr_create(ss.str().c_str(), &ctx);
    if (rc == ZAMR_OK) rc = zamr_init_data(ctx);
    if (rc == ZAMR_OK) rc = zamr_evolve(ctx);
    if (rc != ZAMR_OK) {
        std::fprintf(stderr, "zamr error %d: %s\n", rc, ctx ? zamr_last_error(ctx) : "");
        zamr_destroy(ctx);
        return 1;
    }
    zamr_destroy(ctx);
    return 0;
}
