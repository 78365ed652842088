#include "arma_io.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace tbdhost {

static const char* kHeader = "ARMA_CUB_BIN_FC016";

void WriteCubeCplx(std::ostream& out, const cx* data, long long rows, long long cols, long long slices) {
  out << kHeader << '\n' << rows << ' ' << cols << ' ' << slices << '\n';
  out.write(reinterpret_cast<const char*>(data), (std::streamsize)(sizeof(cx) * rows * cols * slices));
  if (!out) throw std::runtime_error("RealCplxCube::SaveCplx(ofstream,ftype_t) > Write Failed!");
}

void WriteCubeReal(std::ostream& out, const double* data, long long rows, long long cols, long long slices) {
  out << "ARMA_CUB_BIN_FN008" << '\n' << rows << ' ' << cols << ' ' << slices << '\n';
  out.write(reinterpret_cast<const char*>(data), (std::streamsize)(sizeof(double) * rows * cols * slices));
  if (!out) throw std::runtime_error("RealCplxCube::SaveReal(ofstream,ftype_t) > Write Failed!");
}

void ReadCubeCplx(std::istream& in, std::vector<cx>& data, long long rows, long long cols, long long slices) {
  std::string header;
  std::getline(in, header);
  if (header != kHeader) throw std::runtime_error("RealCplxCube::LoadCplx(ifstream,ftype_t) > Unexpected Header '" + header + "'");
  long long r = 0, c = 0, s = 0;
  std::string dims;
  std::getline(in, dims);
  std::istringstream ds(dims);
  ds >> r >> c >> s;
  if (r != rows || c != cols || s != slices) {
    std::ostringstream m;
    m << "RealCplxCube::LoadCplx(ifstream,ftype_t) > Size Mismatch: file " << r << "x" << c << "x" << s
      << ", expected " << rows << "x" << cols << "x" << slices;
    throw std::runtime_error(m.str());
  }
  data.resize((size_t)(rows * cols * slices));
  in.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(sizeof(cx) * data.size()));
  if (!in) throw std::runtime_error("RealCplxCube::LoadCplx(ifstream,ftype_t) > Truncated Data!");
}

static std::ifstream OpenIn(const std::string& path) {
  std::ifstream in(path, std::ios::in | std::ios::binary);
  if (!in.is_open()) throw std::runtime_error("Variable::Load(string,ioctrl_t) > Can not Open the Input File! (" + path + ")");
  std::string level;
  std::getline(in, level);
  if (level != "IOControlLevelOne")
    throw std::runtime_error("Variable::Load(int,ioctrl_t,string) > Wrong Request of IO Level One!");
  return in;
}

void SaveCarrierVariable(const std::string& path, const cx* uvw, const cx* p, int nz, int nyc, int mx) {
  std::ofstream out(path, std::ios::out | std::ios::binary);
  if (!out.is_open()) throw std::runtime_error("FlatCarrierVariable::Save(int,ioctrl_t,string) > Can not Open " + path);
  out << "IOControlLevelOne" << std::endl;
  WriteCubeCplx(out, uvw, nz, nyc, 3LL * mx);
  WriteCubeCplx(out, p, nz, nyc, mx);
}

void LoadCarrierVariable(const std::string& path, std::vector<cx>& uvw, std::vector<cx>& p, int nz, int nyc, int mx) {
  std::ifstream in = OpenIn(path);
  ReadCubeCplx(in, uvw, nz, nyc, 3LL * mx);
  ReadCubeCplx(in, p, nz, nyc, mx);
}

void SaveScalarVariable(const std::string& path, const cx* c, int nz, int nyc, int mx) {
  std::ofstream out(path, std::ios::out | std::ios::binary);
  if (!out.is_open()) throw std::runtime_error("FlatScalarVariable::Save(int,ioctrl_t,string) > Can not Open " + path);
  out << "IOControlLevelOne" << std::endl;
  WriteCubeCplx(out, c, nz, nyc, mx);
}

void LoadScalarVariable(const std::string& path, std::vector<cx>& c, int nz, int nyc, int mx) {
  std::ifstream in = OpenIn(path);
  ReadCubeCplx(in, c, nz, nyc, mx);
}

}  // namespace tbdhost
