#include "baseUtilityLib.h"

#include <random>

using namespace baseUtl;

Matrix::Matrix(int m, int n, int ld, const float *matrix) : m(m), n(n), ld(ld)
{
    hostVector = new std::vector<float>(matrix, matrix + (size_t)m * n);
}

Matrix::Matrix(int m, int n, int ld, std::vector<float> *matrix) : m(m), n(n), ld(ld), hostVector(matrix) {}

Matrix::~Matrix() { delete hostVector; }

Matrix *Matrix::cloneHost()
{
    return new Matrix(m, n, ld, new std::vector<float>(*hostVector));
}

Matrix *Matrix::randomMatrixHost(int m, int n, int ld, uint64_t seed)
{
    std::mt19937_64 engine(seed);
    std::vector<float> *v = new std::vector<float>((size_t)ld * n);
    for (auto &x : *v) x = (float)((engine() >> 11) * 0x1.0p-53);
    return new Matrix(m, n, ld, v);
}

int64_t TimeElapsed::getTotalTime() { return init + working + finalize; }
