// baseUtl: the reference's column-major Matrix and phase timers
// (src/lib/baseUtilityLib/baseUtilityLib.h:16-44), kept with the same public
// fields.  Device memory is owned by the HIP context (ksvd_ctx), so a Matrix
// here only carries host data; it owns its vector (the reference leaks and
// deletes hostVector->data(), matrix.cu:72-82).
#pragma once

#include <cstdint>
#include <vector>

namespace baseUtl {

class Matrix;
class TimeElapsed;

class Matrix {
public:
    int m, n, ld;
    std::vector<float> *hostVector = nullptr;

    // Copy of m*n floats (matrix.cu:13-19).
    Matrix(int m, int n, int ld, const float *matrix);
    // Takes ownership of the vector (matrix.cu:28-34).
    Matrix(int m, int n, int ld, std::vector<float> *matrix);
    ~Matrix();
    Matrix(const Matrix &) = delete;
    Matrix &operator=(const Matrix &) = delete;

    Matrix *cloneHost();
    // Uniform [0,1) entries (matrix.cu:91-105), from a seeded generator.
    static Matrix *randomMatrixHost(int m, int n, int ld, uint64_t seed = 1);
};

class TimeElapsed {
public:
    int64_t init = 0, working = 0, finalize = 0;  // milliseconds
    int64_t getTotalTime();
};

}  // namespace baseUtl
