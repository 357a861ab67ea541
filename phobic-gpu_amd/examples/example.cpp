// example.cpp -- build an MPHF over string keys on the MI355X and check it is a
// bijection onto [0, n) (the validity contract of the reference's
// example.cpp:43-59).  Usage: example [n] [lambda] [P] [small]
// ("small" builds a SmallSpaceMphf instead of a FastQueryMphf).
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include <phobic_gpu_mphf.hpp>

using namespace phobicgpu;

template <typename Mphf>
int run(size_t size, double lambda, uint32_t P) {
  std::vector<std::string> keys;
  keys.reserve(size);
  for (size_t i = 0; i < size; ++i) keys.push_back(std::to_string(i));

  MPHFbuilder builder(MPHFconfig(lambda, P));
  Mphf f;
  auto begin = std::chrono::high_resolution_clock::now();
  HostTimer t = builder.build(keys, f);
  auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::high_resolution_clock::now() - begin);
  std::cout << "Construction took " << ms.count() << " ms" << std::endl;
  std::cout << "RESULT " << f.getResultLine() << " " << t.getResultStyle(double(size)) << "size=" << size
            << " l=" << lambda << " partition_size=" << P << std::endl;

  std::vector<bool> taken(keys.size(), false);
  for (size_t i = 0; i < keys.size(); i++) {
    const size_t h = f(keys[i]);
    if (h >= keys.size()) {
      std::cerr << "Out of range!" << std::endl;
      return 1;
    }
    if (taken[h]) {
      std::cerr << "Collision by key " << i << "!" << std::endl;
      return 1;
    }
    taken[h] = true;
  }
  std::cout << "valid for " << keys.size() << std::endl;
  return 0;
}

int main(int argc, char** argv) {
  const size_t size = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 10000000;
  const double lambda = argc > 2 ? std::atof(argv[2]) : 4.5;
  const uint32_t P = argc > 3 ? (uint32_t)std::strtoul(argv[3], nullptr, 10) : 2500;
  if (argc > 4 && std::string(argv[4]) == "small") return run<SmallSpaceMphf>(size, lambda, P);
  return run<FastQueryMphf>(size, lambda, P);
}
