// Armadillo binary cube format of the state files (Armadillo's
// arma_binary: header "ARMA_CUB_BIN_FC016", a line "rows cols slices", then
// the raw column-major complex<double> data; reference
// src/cube_real_cplx.cc:820-891, tmtools/w_armadillo_cplx.m:35-40) and the
// phase_NN_variable.dat layout (src/variable_carrier_flat.cc:143-377,
// src/variable_scalar_flat.cc:59-212): a text line "IOControlLevelOne" then
// the cubes.
#pragma once

#include <complex>
#include <istream>
#include <ostream>
#include <string>
#include <vector>

namespace tbdhost {

typedef std::complex<double> cx;

void WriteCubeCplx(std::ostream& out, const cx* data, long long rows, long long cols, long long slices);
// real cube (header "ARMA_CUB_BIN_FN008"; RealCplxCube::SaveReal,
// src/cube_real_cplx.cc:797-809)
void WriteCubeReal(std::ostream& out, const double* data, long long rows, long long cols, long long slices);
// Reads one cube; throws if the header or the expected shape does not match.
void ReadCubeCplx(std::istream& in, std::vector<cx>& data, long long rows, long long cols, long long slices);

// <dir>/phase_NN_variable.dat: carrier (uvw: rows=nz, cols=nyc, slices=3*mx;
// p: slices=mx), scalar (c).
void SaveCarrierVariable(const std::string& path, const cx* uvw, const cx* p, int nz, int nyc, int mx);
void LoadCarrierVariable(const std::string& path, std::vector<cx>& uvw, std::vector<cx>& p, int nz, int nyc, int mx);
void SaveScalarVariable(const std::string& path, const cx* c, int nz, int nyc, int mx);
void LoadScalarVariable(const std::string& path, std::vector<cx>& c, int nz, int nyc, int mx);

}  // namespace tbdhost
