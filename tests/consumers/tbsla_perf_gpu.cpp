// A consumer of the tbsla::cpp C++ API, compiled against include/tbsla/cpp/*.hpp
// and linked to libtbsla_hip.so, used the way the reference's own benchmark
// application uses its library (src/cpp/apps/tbsla_perf.cpp:92-143): the
// matrix is held through the tbsla::cpp::Matrix base and every call is
// virtual.  It is built by __graft_entry__.build() and run by
// tests/test_gpu_consumers.py.
//
//   tbsla_perf_gpu --format CSR|COO|SCOO|ELL|DENSE
//                  --matrix cdiag|cqmat|stochastic|<binary matrix file>
//                  --NR n --NC n [--C c --Q q --S s --pr p --pc p --GR g --GC g]
//                  --op spmv|Ax|a_axpx|page_rank|personalized_page_rank
//                  --x <vector file> --out <vector file>
//                  [--epsilon e --beta b --max_iteration k --nodes i,j,...]
//
// Vector files hold a size_t count then the doubles (the layout of
// C_CPP_Vector_write, interface/c/matrix/Vector.cpp:90-103).  One JSON line
// with the reference application's keys goes to stdout.
#include <tbsla/cpp/Matrix.hpp>
#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::map<std::string, std::string> parse(int argc, char **argv) {
  std::map<std::string, std::string> o;
  for (int i = 1; i + 1 < argc; i += 2) o[argv[i]] = argv[i + 1];
  return o;
}

std::string opt(const std::map<std::string, std::string> &o, const std::string &k, const std::string &d = "") {
  auto it = o.find(k);
  return it == o.end() ? d : it->second;
}

std::vector<double> read_vector(const std::string &path) {
  std::ifstream is(path, std::ios::binary);
  if (!is) throw std::runtime_error("cannot open " + path);
  std::size_t n = 0;
  is.read(reinterpret_cast<char *>(&n), sizeof(n));
  std::vector<double> v(n);
  is.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(double)));
  if (!is) throw std::runtime_error("short vector file " + path);
  return v;
}

void write_vector(const std::string &path, const std::vector<double> &v) {
  std::ofstream os(path, std::ios::binary | std::ios::trunc);
  std::size_t n = v.size();
  os.write(reinterpret_cast<const char *>(&n), sizeof(n));
  os.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(n * sizeof(double)));
}

tbsla::cpp::Matrix *make(const std::string &f) {
  if (f == "COO") return new tbsla::cpp::MatrixCOO();
  if (f == "SCOO") return new tbsla::cpp::MatrixSCOO();
  if (f == "CSR") return new tbsla::cpp::MatrixCSR();
  if (f == "ELL") return new tbsla::cpp::MatrixELL();
  if (f == "DENSE") return new tbsla::cpp::MatrixDENSE();
  return nullptr;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char **argv) {
  const auto o = parse(argc, argv);
  const std::string format = opt(o, "--format", "CSR"), matrix = opt(o, "--matrix"), op = opt(o, "--op", "spmv");
  std::unique_ptr<tbsla::cpp::Matrix> m(make(format));
  if (!m) {
    std::cerr << format << " unrecognized!" << std::endl;
    return 1;
  }
  const int NR = std::stoi(opt(o, "--NR", "0")), NC = std::stoi(opt(o, "--NC", "0"));
  const int C = std::stoi(opt(o, "--C", "8"));
  const double Q = std::stod(opt(o, "--Q", "0.1"));
  const unsigned S = static_cast<unsigned>(std::stoul(opt(o, "--S", "0")));
  const int pr = std::stoi(opt(o, "--pr", "0")), pc = std::stoi(opt(o, "--pc", "0"));
  const int GR = std::stoi(opt(o, "--GR", "1")), GC = std::stoi(opt(o, "--GC", "1"));
  try {
    if (matrix == "cdiag") {
      m->fill_cdiag(NR, NC, C, pr, pc, GR, GC);
    } else if (matrix == "cqmat") {
      m->fill_cqmat(NR, NC, C, Q, S, pr, pc, GR, GC);
    } else if (matrix == "stochastic") {
      m->fill_cqmat_stochastic(NR, NC, C, Q, S, pr, pc, GR, GC);
    } else {
      std::ifstream is(matrix, std::ifstream::binary);
      if (!is) throw std::runtime_error("cannot open " + matrix);
      m->read(is);
    }
    std::vector<double> result;
    int iterations = -1;
    const double t0 = now();
    if (op == "spmv") {
      result = m->spmv(read_vector(opt(o, "--x")));
    } else if (op == "Ax") {
      result.assign(static_cast<std::size_t>(m->get_n_row()), 0.0);
      m->Ax(result, read_vector(opt(o, "--x")));
    } else if (op == "a_axpx") {
      result = m->a_axpx_(read_vector(opt(o, "--x")));
    } else if (op == "page_rank") {
      result = m->page_rank(std::stod(opt(o, "--epsilon", "1e-10")), std::stod(opt(o, "--beta", "0.85")),
                            std::stoi(opt(o, "--max_iteration", "100")), iterations);
    } else if (op == "personalized_page_rank") {
      std::vector<int> nodes;
      std::stringstream ss(opt(o, "--nodes"));
      for (std::string tok; std::getline(ss, tok, ',');) nodes.push_back(std::stoi(tok));
      result = m->personalized_page_rank(std::stod(opt(o, "--epsilon", "1e-10")), std::stod(opt(o, "--beta", "0.85")),
                                         std::stoi(opt(o, "--max_iteration", "100")), nodes, iterations);
    } else {
      std::cerr << "OP : " << op << " unrecognized!" << std::endl;
      return 1;
    }
    const double t1 = now();
    write_vector(opt(o, "--out"), result);
    const double flops = (op == "a_axpx" ? 4.0 : 2.0) * static_cast<double>(m->get_nnz());
    std::cout << "{\"test\":\"" << op << "\",\"matrix_format\":\"" << format << "\",\"n_row\":" << m->get_n_row()
              << ",\"n_col\":" << m->get_n_col() << ",\"nnz\":" << m->get_nnz() << ",\"time_op\":" << (t1 - t0)
              << ",\"gflops\":" << flops / (t1 - t0) / 1e9 << ",\"iterations\":" << iterations
              << ",\"lang\":\"CPP_HIP\"}" << std::endl;
  } catch (const std::exception &e) {
    std::cerr << "tbsla_perf_gpu: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
