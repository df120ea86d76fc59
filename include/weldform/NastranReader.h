// weldform/NastranReader.h -- the reference's reader of Nastran bulk-data
// triangle meshes (Source/NastranReader.h:26-204), the source of
// TriMesh(NastranReader &, flipnormals) rigid bodies ("RigidBodies" of type
// "File" in a scene).  Fixed 8-character fields; GRID cards carry the node
// id in field 2 and x, y, z in fields 4-6, CTRIA* cards the three node ids
// in fields 4-6.  As in the reference the GRID cards are read as one
// contiguous run from the first of them, and so are the element cards;
// numbers may use Nastran's exponent without the E ("0.5633-18",
// "-.962-03"), which the reader inserts before the first '+', or before the
// first '-' that is not the leading sign.  CBEAM / CBAR (2D) meshes are
// counted but rejected: the device path is 3D.
#pragma once
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "matvec.h"

namespace SPH {

class NastranReader {
 public:
  int dim = 3;
  int node_count = 0, elem_count = 0;
  std::vector<double> node;     // 3 per node
  std::vector<int> elcon;       // 3 node positions per element
  std::vector<int> nodeid;      // the file's node ids
  std::map<int, int> nodepos;   // id -> position

  NastranReader() {}
  explicit NastranReader(const char *fName) { read(fName); }

  // a field of a fixed-format card: the reference's substr(pos, 8)
  static std::string field(const std::string &line, size_t pos) {
    if (pos > line.size()) throw new Fatal("NastranReader: card too short: " + line);
    return line.substr(pos, 8);
  }
  // a real number in Nastran's short form
  static double real(std::string t) {
    size_t p = t.find('+');
    if (p != std::string::npos) {
      t.insert(p, "E");
    } else {
      p = t.find('-');
      if (p != std::string::npos) {
        if (p != 0) {
          t.insert(p, "E");
        } else {
          p = t.find('-', 1);
          if (p != std::string::npos) t.insert(p, "E");
        }
      }
    }
    return std::strtod(t.c_str(), nullptr);
  }

  void read(const char *fName) {
    std::ifstream f(fName);
    if (!f) throw new Fatal(std::string("NastranReader: cannot open ") + fName);
    std::vector<std::string> raw;
    std::string line;
    int first_node = -1, first_elem = -1;
    dim = 3;
    node_count = elem_count = 0;
    while (std::getline(f, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      const int l = (int)raw.size();
      raw.push_back(line);
      if (line.compare(0, 4, "GRID") == 0) {
        if (first_node < 0) first_node = l;
        ++node_count;
      } else if (line.compare(0, 5, "CTRIA") == 0 || line.compare(0, 5, "CBEAM") == 0 ||
                 line.compare(0, 4, "CBAR") == 0) {
        if (first_elem < 0) first_elem = l;
        if (line.compare(0, 5, "CTRIA") != 0) dim = 2;
        ++elem_count;
      }
    }
    if (node_count == 0 || elem_count == 0) throw new Fatal(std::string("NastranReader: no GRID or element cards in ") + fName);
    if (dim != 3) throw new Fatal("NastranReader: 2D (CBEAM / CBAR) meshes are not supported on the device path");
    node.assign(3 * (size_t)node_count, 0.0);
    nodeid.assign(node_count, 0);
    nodepos.clear();
    for (int n = 0; n < node_count; ++n) {
      const std::string &c = raw[first_node + n];
      if (c.compare(0, 4, "GRID") != 0) throw new Fatal("NastranReader: GRID cards are not contiguous");
      nodeid[n] = std::atoi(field(c, 8).c_str());
      nodepos.insert(std::make_pair(nodeid[n], n));
      for (int i = 0; i < 3; ++i) node[3 * (size_t)n + i] = real(field(c, 24 + 8 * i));
    }
    elcon.assign(3 * (size_t)elem_count, 0);
    for (int e = 0; e < elem_count; ++e) {
      const std::string &c = raw[first_elem + e];
      if (c.compare(0, 5, "CTRIA") != 0) throw new Fatal("NastranReader: element cards are not contiguous");
      for (int k = 0; k < 3; ++k) {
        const int id = std::atoi(field(c, 24 + 8 * k).c_str());
        auto it = nodepos.find(id);
        if (it == nodepos.end()) throw new Fatal("NastranReader: element references an unknown node id");
        elcon[3 * (size_t)e + k] = it->second;
      }
    }
  }
};

}  // namespace SPH
