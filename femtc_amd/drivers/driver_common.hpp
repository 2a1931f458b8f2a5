// Shared body of the bench_mass / bench_stiffness drivers: the reference's
// command line (<Ndofs> <degree> <order> <Mb> <Nb> <type>, bench_mass.cpp:
// 170-185) and result line (degree nranks type Ndofs elapsed order Mb Nb
// GFLOP/s GB/s GDoF/s, bench_mass.cpp:142-145), with the operator applied on
// the GPU through the femtc C ABI.  Data is resident on the device during the
// timed applies; as in the reference, 10 applies are run and the mean of the
// last 5 is reported (bench_mass.cpp:95-106).  One process = one "rank".
//
// Optional 7th argument: "reference" (default; the reference's arithmetic)
// or "correct" (the mathematically intended operator).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "femtc.h"
#include "femtc/linalg.hpp"

namespace driver
{
inline void check(int st, const char* what)
{
  if (st != FEMTC_OK)
    throw std::runtime_error(std::string(what) + ": " +
                             femtc_status_string(st) + " (" +
                             femtc_last_error() + ")");
}

struct DeviceArray
{
  void* p = nullptr;
  explicit DeviceArray(std::size_t bytes)
  {
    check(femtc_device_alloc(&p, bytes), "device_alloc");
  }
  ~DeviceArray() { femtc_device_free(p); }
};

template <typename T>
std::vector<T> random_vector(std::size_t n)
{
  std::vector<T> v(n);
  for (auto& x : v)
    x = std::rand() / (T)RAND_MAX;
  return v;
}

template <typename T>
const char* type_name()
{
  return sizeof(T) == 8 ? "double" : "float";
}

// is_stiffness selects the operator; geometry has 1 (detJ) or 6 (G) values
// per point.  Returns the process exit code.
template <typename T>
int run(bool is_stiffness, int degree, int Ndofs, const std::string& order_str,
        int Mb, int Nb, int mode)
{
  const linalg::Order order = linalg::string2order(order_str);
  const int p = degree;
  const int ndofs = (p + 1) * (p + 1) * (p + 1);
  const int num_cells = Ndofs / ndofs;
  if (num_cells == 0)
    throw std::runtime_error("Number of cells is 0");
  if (degree < 1)
    throw std::runtime_error("Degree must be at least 1");
  const int st_ladder = femtc_check_ladder(degree, (int)order, Mb, Nb);
  if (st_ladder != FEMTC_OK)
  {
    std::cout << femtc_status_string(st_ladder) << std::endl;
    return 0;
  }
  const int N = p + 1;
  const int ntab = (is_stiffness ? 2 : 1) * N * N;
  const int gsize = is_stiffness ? 6 : 1;
  const int dtype = sizeof(T) == 8 ? FEMTC_F64 : FEMTC_F32;

  std::vector<T> phi = random_vector<T>(ntab);
  std::vector<T> U = random_vector<T>((std::size_t)num_cells * ndofs);
  std::vector<T> W((std::size_t)num_cells * ndofs, 0);
  std::vector<T> geom = random_vector<T>((std::size_t)num_cells * ndofs * gsize);

  DeviceArray dphi(phi.size() * sizeof(T)), dU(U.size() * sizeof(T)),
      dW(W.size() * sizeof(T)), dgeom(geom.size() * sizeof(T));
  check(femtc_memcpy_htod(dphi.p, phi.data(), phi.size() * sizeof(T)), "copy");
  check(femtc_memcpy_htod(dU.p, U.data(), U.size() * sizeof(T)), "copy");
  check(femtc_memcpy_htod(dgeom.p, geom.data(), geom.size() * sizeof(T)),
        "copy");

  double elapsed = 0.0;
  for (int i = 0; i < 10; i++)
  {
    check(femtc_device_synchronize(), "synchronize");
    const auto t0 = std::chrono::steady_clock::now();
    const int st =
        is_stiffness
            ? femtc_stiffness_apply(dtype, degree, N, mode, dphi.p, dU.p,
                                    dW.p, dgeom.p, num_cells, nullptr)
            : femtc_mass_apply(dtype, degree, N, mode, dphi.p, dU.p, dW.p,
                               dgeom.p, num_cells, nullptr);
    check(st, "apply");
    check(femtc_device_synchronize(), "synchronize");
    const auto t1 = std::chrono::steady_clock::now();
    if (i >= 5)
      elapsed += std::chrono::duration<double>(t1 - t0).count();
  }
  elapsed /= 5.0;
  check(femtc_memcpy_dtoh(W.data(), dW.p, W.size() * sizeof(T)), "copy");
  for (std::size_t i = 0; i < W.size(); i++)
    if (W[i] < 0)
      throw std::runtime_error("W is negative");

  // the reference's models (bench_mass.cpp:118-129; x3 flops for stiffness,
  // bench_stiffness.cpp:122)
  const double m = N, n = (double)N * N, k = N;
  const double flops =
      (is_stiffness ? 3.0 : 1.0) * 3 * 2 * (2.0 * m * n * k + n * k) * num_cells;
  const double mem = (U.size() + 2.0 * W.size() + geom.size()) * sizeof(T);
  std::cout << degree << " " << 1 << " " << type_name<T>() << " ";
  std::cout << Ndofs << " " << elapsed << " " << order_str << " ";
  std::cout << Mb << " " << Nb << " ";
  std::cout << flops / elapsed / 1e9 << " " << mem / elapsed / 1e9 << " "
            << (double)num_cells * ndofs / elapsed / 1e9 << std::endl;
  return 0;
}

inline int main_impl(bool is_stiffness, int argc, char** argv)
{
  if (argc <= 6)
  {
    std::cout << "Usage: " << argv[0]
              << " <Ndofs> <degree> <order> <Mb> <Nb> <type> [reference|correct]"
              << std::endl;
    return 1;
  }
  const int Ndofs = std::atoi(argv[1]);
  const int degree = std::atoi(argv[2]);
  const std::string order_str = argv[3];
  const int Mb = std::atoi(argv[4]);
  const int Nb = std::atoi(argv[5]);
  const std::string type_str = argv[6];
  int mode = FEMTC_MODE_REFERENCE;
  if (argc > 7 && std::string(argv[7]) == "correct")
    mode = FEMTC_MODE_CORRECT;
  if (type_str == "float")
    return run<float>(is_stiffness, degree, Ndofs, order_str, Mb, Nb, mode);
  return run<double>(is_stiffness, degree, Ndofs, order_str, Mb, Nb, mode);
}
} // namespace driver
