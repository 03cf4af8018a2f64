#include "screenutils.h"

#include <iomanip>
#include <iostream>

void ScreenUtils::PrintScrStarLine(void) { PrintScrCharLine('*'); }

void ScreenUtils::PrintScrCharLine(char h) {
  for (int i = 0; i < 80; ++i) std::cout << h;
  std::cout << std::endl;
}

void ScreenUtils::PrintVectorN(size_t n, const double* v) {
  for (size_t i = 0; i < n; ++i) std::cout << std::setw(12) << v[i];
  std::cout << std::endl;
}

void ScreenUtils::PrintMatrixNxN(size_t n, const double* m) {
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < n; ++j) std::cout << std::setw(12) << m[j * n + i];
    std::cout << std::endl;
  }
}

void ScreenUtils::PrintMatrixNxNSymmetric(size_t n, const double* p) {
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < n; ++j) {
      const size_t r = i > j ? i : j, c = i > j ? j : i;
      std::cout << std::setw(12) << p[r * (r + 1) / 2 + c];
    }
    std::cout << std::endl;
  }
}
