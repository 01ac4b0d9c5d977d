// Binary I/O helpers shared by the reference harnesses.
//
// A harness reads one input file of little-endian float64 values (parameters
// first, then arrays) and writes one output file of float64 values.  The Python
// side (tests/golden/make_golden.py) owns both layouts.  These harnesses are
// test infrastructure: they drive the reference's own CPU classes, compiled in
// place from the reference sources, to produce golden vectors.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <vector>

struct HarnessIn {
    std::vector<double> v;
    size_t pos = 0;
    explicit HarnessIn(const char* path) {
        FILE* f = std::fopen(path, "rb");
        if (!f) { std::perror(path); std::exit(2); }
        double x;
        while (std::fread(&x, sizeof(double), 1, f) == 1) v.push_back(x);
        std::fclose(f);
    }
    double d() {
        if (pos >= v.size()) { std::fprintf(stderr, "harness input exhausted\n"); std::exit(3); }
        return v[pos++];
    }
    int i() { return (int)d(); }
    void arr(double* dst, size_t n) { for (size_t k = 0; k < n; ++k) dst[k] = d(); }
};

struct HarnessOut {
    FILE* f;
    explicit HarnessOut(const char* path) {
        f = std::fopen(path, "wb");
        if (!f) { std::perror(path); std::exit(2); }
    }
    ~HarnessOut() { std::fclose(f); }
    void d(double x) { std::fwrite(&x, sizeof(double), 1, f); }
    void arr(const double* src, size_t n) { std::fwrite(src, sizeof(double), n, f); }
};
