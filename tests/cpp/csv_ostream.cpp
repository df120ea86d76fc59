// Writes the CSV of Domain::WriteCSV (Source/Output.cpp:521-554) from a
// binary table of 17 doubles per row, printing every value through
// std::ostream's default formatting as the reference does (ID as an int).
// The output test compares it byte for byte with wfio_write_csv.
//   csv_ostream rows.bin n out.csv
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <vector>

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  const long n = std::atol(argv[2]);
  std::vector<double> r((size_t)n * 17);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(r.data(), sizeof(double), r.size(), f) != r.size()) return 3;
  std::fclose(f);
  std::ostringstream oss;
  oss << "X, Y, Z, ID, Sigma_eq, Pl_Strain, vx, vy, vz, ax, ay, az, CFx, CFy, CZFz, p, Temp" << std::endl;
  for (long i = 0; i < n; ++i) {
    const double *p = &r[(size_t)i * 17];
    for (int j = 0; j < 3; ++j) oss << p[j] << ", ";
    oss << (int)p[3];
    for (int c = 4; c < 17; ++c) oss << ", " << p[c];
    oss << std::endl;
  }
  std::ofstream of(argv[3], std::ios::out);
  of << oss.str();
  return of ? 0 : 4;
}
