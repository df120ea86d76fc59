// wf_io.cpp -- libweldform_io.so: the reference's output files from the
// packed particle tables (include/weldform_io.h).
//
// HDF5 part: Domain::WriteXDMF (Source/Output.cpp:75-518) writes one 1-D
// dataset per field with the HDF5 lite API and an XDMF 2.0 file naming
// them; the dataset names, their creation order, the element and attribute
// text of the .xmf are kept as the reference writes them, so the files load
// in the same ParaView / XDMF pipelines.  CSV part: Domain::WriteCSV
// (Source/Output.cpp:521-554).
#include "../../include/weldform_io.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/weldform_hip.h"

#ifdef WFIO_HDF5
#include <hdf5.h>
#include <hdf5_hl.h>
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

enum Kind { kScalar, kVector, kTensor6 };

struct Dataset {
  const char *name;  // dataset name (Property1..3 replaced by OutputName)
  int off;           // WF_XDMF_* offset (float) or int slot
  int width;
  bool is_int;
};

// creation order of the datasets in the .hdf5 (Source/Output.cpp:225-294)
const Dataset kCreate[] = {
    {"Position", WF_XDMF_POSITION, 3, false},
    {"Velocity", WF_XDMF_VELOCITY, 3, false},
    {"Acceleration", WF_XDMF_ACCELERATION, 3, false},
    {"Tag", WF_XDMF_TAG, 1, true},
    {"Pressure", WF_XDMF_PRESSURE, 1, false},
    {"Density", WF_XDMF_DENSITY, 1, false},
    {nullptr, WF_XDMF_PROPERTY1, 1, false},
    {nullptr, WF_XDMF_PROPERTY2, 1, false},
    {nullptr, WF_XDMF_PROPERTY3, 1, false},
    {"Mass", WF_XDMF_MASS, 1, false},
    {"h", WF_XDMF_H, 1, false},
    {"Sigma", WF_XDMF_SIGMA, 6, false},
    {"ShearS", WF_XDMF_SHEARS, 6, false},
    {"Strain", WF_XDMF_STRAIN, 6, false},
    {"StrainRate", WF_XDMF_STRAIN_RATE, 6, false},
    {"Strain_pl", WF_XDMF_STRAIN_PL, 6, false},
    {"gradcorrmat", WF_XDMF_GRADCORRMAT, 6, false},
    {"Sigma_eq", WF_XDMF_SIGMA_EQ, 1, false},
    {"Temperature", WF_XDMF_TEMPERATURE, 1, false},
    {"Pl_Strain", WF_XDMF_PL_STRAIN, 1, false},
    {"Neighbors", WF_XDMF_NEIGHBORS, 1, true},
    {"ContNeib", WF_XDMF_CONTNEIB, 1, true},
    {"q_friction", WF_XDMF_Q_FRICTION, 1, false},
    {"c_shearabs", WF_XDMF_C_SHEARABS, 1, false},
    {"ps_en", WF_XDMF_PS_EN, 1, false},
    {"Displacement", WF_XDMF_DISPLACEMENT, 3, false},
    {"Contact Force", WF_XDMF_CONTACT_FORCE, 3, false},
};

// order of the Attribute elements in the .xmf (Source/Output.cpp:352-498);
// names as in kCreate, nullptr k = Property k
struct Attr {
  const char *name;
  int prop;  // 0, or 1..3 for the OutputName datasets
  Kind kind;
  bool is_int;
};
const Attr kAttrs[] = {
    {"Tag", 0, kScalar, true},           {"Position", 0, kVector, false},
    {"Velocity", 0, kVector, false},     {"Acceleration", 0, kVector, false},
    {"Density", 0, kScalar, false},      {"h", 0, kScalar, false},
    {"Pressure", 0, kScalar, false},     {nullptr, 1, kScalar, false},
    {nullptr, 2, kScalar, false},        {nullptr, 3, kScalar, false},
    {"Sigma", 0, kTensor6, false},       {"ShearS", 0, kTensor6, false},
    {"Strain", 0, kTensor6, false},      {"StrainRate", 0, kTensor6, false},
    {"Strain_pl", 0, kTensor6, false},   {"gradcorrmat", 0, kTensor6, false},
    {"Temperature", 0, kScalar, false},  {"Sigma_eq", 0, kScalar, false},
    {"Pl_Strain", 0, kScalar, false},    {"Neighbors", 0, kScalar, true},
    {"ContNeib", 0, kScalar, true},      {"Displacement", 0, kVector, false},
    {"Contact Force", 0, kVector, false}, {"q_friction", 0, kScalar, false},
    {"c_shearabs", 0, kScalar, false},   {"ps_en", 0, kScalar, false},
};

// the .xmf text (element and attribute layout of Source/Output.cpp:340-510)
std::string xmf_text(const std::string &h5, int64_t n, const std::string prop[3]) {
  std::ostringstream o;
  o << "<?xml version=\"1.0\" ?>\n"
    << "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
    << "<Xdmf Version=\"2.0\">\n"
    << " <Domain>\n"
    << "   <Grid Name=\"SPHCenter\" GridType=\"Uniform\">\n"
    << "     <Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" << n << "\"/>\n"
    << "     <Geometry GeometryType=\"XYZ\">\n"
    << "       <DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"10\" Dimensions=\"" << n << " 3\" >\n"
    << "        " << h5 << ":/Position \n"
    << "       </DataItem>\n"
    << "     </Geometry>\n";
  for (const Attr &a : kAttrs) {
    const std::string name = a.prop ? prop[a.prop - 1] : std::string(a.name);
    const char *type = a.kind == kScalar ? "Scalar" : (a.kind == kVector ? "Vector" : "Tensor6");
    o << "     <Attribute Name=\"" << name << "\" AttributeType=\"" << type << "\" Center=\"Node\">\n";
    o << "       <DataItem Dimensions=\"" << n;
    if (a.kind == kVector) o << " 3";
    if (a.kind == kTensor6) o << " 6";
    o << "\" NumberType=\"" << (a.is_int ? "Int" : "Float") << "\"";
    if (a.is_int && name == "Tag")
      o << " Format=\"HDF\">\n";           // Tag carries no Precision
    else if (a.kind == kScalar)
      o << " Precision=\"10\"  Format=\"HDF\">\n";
    else
      o << " Precision=\"10\" Format=\"HDF\">\n";
    o << "        " << h5 << ":/" << name << " \n";
    o << "       </DataItem>\n";
    o << "     </Attribute>\n";
  }
  o << "   </Grid>\n"
    << " </Domain>\n"
    << "</Xdmf>\n";
  return o.str();
}

}  // namespace

extern "C" {

int wfio_has_hdf5(void) {
#ifdef WFIO_HDF5
  return 1;
#else
  return 0;
#endif
}

const char *wfio_last_error(void) { return g_err.c_str(); }

int wfio_write_xdmf(const char *file_key, int64_t n, const float *f, const int32_t *iv,
                    const char *const prop_names[3]) {
  g_err.clear();
  if (!file_key || n < 0 || (n > 0 && (!f || !iv))) return fail(WFIO_EINVAL, "bad argument");
  std::string prop[3] = {"Property1", "Property2", "Property3"};  // Domain::OutputName defaults
  if (prop_names)
    for (int k = 0; k < 3; ++k)
      if (prop_names[k]) prop[k] = prop_names[k];
#ifndef WFIO_HDF5
  (void)prop;
  return fail(WFIO_ENOHDF5, "libweldform_io was built without HDF5: XDMF output is unavailable");
#else
  const std::string h5 = std::string(file_key) + ".hdf5";
  const hid_t file = H5Fcreate(h5.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return fail(WFIO_EIO, "cannot create " + h5);
  bool ok = true;
  {
    const int np[1] = {(int)n};
    const hsize_t one[1] = {1};
    ok = ok && H5LTmake_dataset_int(file, "/NP", 1, one, np) >= 0;
  }
  for (const Dataset &d : kCreate) {
    if (!ok) break;
    std::string name = d.name ? d.name : prop[d.off - WF_XDMF_PROPERTY1];
    const hsize_t dims[1] = {(hsize_t)(d.width * n)};
    if (d.is_int)
      ok = H5LTmake_dataset_int(file, name.c_str(), 1, dims, (const int *)iv + (size_t)d.off * n) >= 0;
    else
      ok = H5LTmake_dataset_float(file, name.c_str(), 1, dims, f + (size_t)d.off * n) >= 0;
    if (!ok) g_err = "cannot write dataset " + name + " of " + h5;
  }
  if (H5Fflush(file, H5F_SCOPE_GLOBAL) < 0) ok = false;
  if (H5Fclose(file) < 0) ok = false;
  if (!ok) return fail(WFIO_EIO, g_err.empty() ? "HDF5 error writing " + h5 : g_err);
  const std::string xmf = std::string(file_key) + ".xmf";
  std::ofstream of(xmf, std::ios::out);
  of << xmf_text(h5, n, prop);
  of.close();
  if (!of) return fail(WFIO_EIO, "cannot write " + xmf);
  return WFIO_OK;
#endif
}

int wfio_write_csv(const char *file_key, int64_t n, const double *rows, int32_t threads) {
  if (!file_key || n < 0 || (n > 0 && !rows) || threads < 1) return fail(WFIO_EINVAL, "bad argument");
  const int nt = (int)std::min<int64_t>(threads, n > 0 ? n : 1);
  std::vector<std::string> part((size_t)nt);
  auto format = [&](int t) {
    const int64_t a = n * t / nt, b = n * (t + 1) / nt;
    std::string &s = part[(size_t)t];
    s.reserve((size_t)(b - a) * 200);
    char buf[64];
    for (int64_t i = a; i < b; ++i) {
      const double *r = rows + (size_t)i * WF_CSV_NCOL;
      for (int c = 0; c < WF_CSV_NCOL; ++c) {
        int len;
        if (c == 3)
          len = std::snprintf(buf, sizeof buf, "%lld", (long long)r[c]);
        else
          len = std::snprintf(buf, sizeof buf, "%g", r[c]);
        s.append(buf, (size_t)len);
        if (c + 1 < WF_CSV_NCOL) s.append(", ");
      }
      s.push_back('\n');
    }
  };
  if (nt == 1) {
    format(0);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(format, t);
    for (auto &th : pool) th.join();
  }
  const std::string fn = std::string(file_key) + ".csv";
  std::ofstream of(fn, std::ios::out | std::ios::binary);
  of << "X, Y, Z, ID, Sigma_eq, Pl_Strain, vx, vy, vz, ax, ay, az, CFx, CFy, CZFz, p, Temp\n";
  for (const std::string &s : part) of.write(s.data(), (std::streamsize)s.size());
  of.close();
  if (!of) return fail(WFIO_EIO, "cannot write " + fn);
  return WFIO_OK;
}

}  // extern "C"
