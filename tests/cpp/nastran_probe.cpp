// Reads a Nastran mesh with the facade's NastranReader and TriMesh(reader,
// flip) and prints, one per line: the node coordinates (%.17g), then each
// element's node positions, centroid and normal.
//   nastran_probe mesh.nas flip
#include <cstdio>
#include <cstdlib>

#include "weldform/Mesh.h"

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  try {
    SPH::NastranReader nr(argv[1]);
    SPH::TriMesh M(nr, std::atoi(argv[2]) != 0);
    std::printf("%d %d\n", nr.node_count, nr.elem_count);
    for (int n = 0; n < nr.node_count; ++n)
      std::printf("%.17g %.17g %.17g\n", nr.node[3 * n], nr.node[3 * n + 1], nr.node[3 * n + 2]);
    for (size_t e = 0; e < M.element.Size(); ++e) {
      const SPH::Element *el = M.element[e];
      std::printf("%d %d %d %.17g %.17g %.17g %.17g %.17g %.17g\n", el->node[0], el->node[1], el->node[2],
                  el->centroid(0), el->centroid(1), el->centroid(2), el->normal(0), el->normal(1), el->normal(2));
    }
  } catch (Fatal *f) {
    std::fprintf(stderr, "%s\n", f->Msg().c_str());
    return 1;
  }
  return 0;
}
