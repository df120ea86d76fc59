// weldform/Mesh.h -- rigid triangle meshes for contact, with the interface
// of the reference's TriMesh (Source/Mesh.h:30-90): AxisPlaneMesh builds a
// plane of 2 dens^2 triangles (Source/Mesh.cpp:82-202), TriMesh(reader,
// flip) a mesh read from a Nastran file (:17-64), CalcSpheres the element
// radii and plane coefficients (:206-227).  The device receives a plane's
// parameters, or a file mesh's nodes and triangles, and rebuilds the same
// mesh; this host copy supplies the particle positions of
// Domain::AddTrimeshParticles.
#pragma once
#include <cmath>

#include "NastranReader.h"
#include "matvec.h"

namespace SPH {

class Element {
 public:
  Element() {}
  Element(const int &n1, const int &n2, const int &n3) : node{n1, n2, n3} {}
  int node[3] = {0, 0, 0};
  Vec3_t centroid, normal;
  double pplane = 0.0, radius = 0.0;
  int nfar = 0;
};

class TriMesh {
 public:
  TriMesh() {}
  ~TriMesh() {
    for (size_t i = 0; i < node.Size(); ++i) delete node[i];
    for (size_t i = 0; i < node_v.Size(); ++i) delete node_v[i];
    for (size_t i = 0; i < element.Size(); ++i) delete element[i];
  }
  TriMesh(const TriMesh &) = delete;
  TriMesh &operator=(const TriMesh &) = delete;

  Array<Vec3_t *> node, node_v;
  Array<Element *> element;
  Vec3_t m_v, m_w;
  int id = 0;
  int dimension = 3;
  // the plane's parameters, passed to the device (dens = 0: a mesh read
  // from a file, passed as nodes and triangles)
  int axis = 2, dens = 0;
  bool flip = false;
  Vec3_t p1, p2;

  // TriMesh(NastranReader &, flipnormals), 3D: nodes as read, triangles
  // with their first two nodes swapped when flipnormals is set; centroids,
  // and normals from the file's winding -- the swap shows only once normals
  // are recomputed (a translation, Scale, or the per-step TriMesh::Update)
  TriMesh(const NastranReader &nr, bool flipnormals) {
    if (nr.dim != 3) throw new Fatal("TriMesh: only 3D meshes are supported");
    for (int n = 0; n < nr.node_count; ++n) {
      node.Push(new Vec3_t(nr.node[3 * n], nr.node[3 * n + 1], nr.node[3 * n + 2]));
      node_v.Push(new Vec3_t(0., 0., 0.));
    }
    for (int e = 0; e < nr.elem_count; ++e) {
      const int a = nr.elcon[3 * e], b = nr.elcon[3 * e + 1], c = nr.elcon[3 * e + 2];
      Element *el = flipnormals ? new Element(b, a, c) : new Element(a, b, c);
      for (int k = 0; k < 3; ++k) el->centroid(k) = ((*node[a])(k) + (*node[b])(k) + (*node[c])(k)) / 3.;
      const Vec3_t w = cross(*node[b] - *node[a], *node[c] - *node[a]);
      el->normal = w / norm(w);
      element.Push(el);
    }
  }

  // TriMesh::CalcCentroids (Source/Mesh.cpp:71-79)
  void CalcCentroids() {
    for (size_t e = 0; e < element.Size(); ++e) {
      Vec3_t c;
      for (int i = 0; i < 3; ++i) c = c + *node[element[e]->node[i]];
      element[e]->centroid = c / 3.;
    }
  }
  // TriMesh::Scale (Source/Mesh.cpp:310-323): nodes scaled, centroids and
  // normals recomputed
  void Scale(const double &f) {
    for (size_t n = 0; n < node.Size(); ++n) *node[n] = f * *node[n];
    CalcCentroids();
    CalcNormals();
  }

  void SetVel(const Vec3_t &v) { m_v = v; }
  void SetRotAxisVel(const Vec3_t &w) { m_w = w; }

  // TriMesh::AxisPlaneMesh, 3D: square elements of side (p2-p1)[dir0]/dens
  // in both in-plane directions, each split into two triangles
  void AxisPlaneMesh(const int &ax, bool positaxisorent, const Vec3_t P1, const Vec3_t P2, const int &ds) {
    if (dimension != 3) throw new Fatal("TriMesh: only 3D plane meshes are supported");
    if (ds < 1) throw new Fatal("TriMesh: dens must be >= 1");
    axis = ax;
    flip = positaxisorent;
    p1 = P1;
    p2 = P2;
    dens = ds;
    const Vec3_t p = P2 - P1;
    int dir[3];
    if (ax == 0) {
      dir[0] = 1;
      dir[1] = 2;
    } else if (ax == 1) {
      dir[0] = 0;
      dir[1] = 2;
    } else {
      dir[0] = 0;
      dir[1] = 1;
    }
    dir[2] = ax;
    const double x3 = P1(dir[2]);
    double x2 = P1(dir[1]);
    const double dl = p(dir[0]) / ds;
    for (int j = 0; j < ds + 1; ++j) {
      double x1 = P1(dir[0]);
      for (int i = 0; i < ds + 1; ++i) {
        Vec3_t v;
        v(dir[0]) = x1;
        v(dir[1]) = x2;
        v(dir[2]) = x3;
        node.Push(new Vec3_t(v));
        node_v.Push(new Vec3_t(0., 0., 0.));
        x1 += dl;
      }
      x2 += dl;
    }
    for (int j = 0; j < ds; ++j)
      for (int i = 0; i < ds; ++i) {
        const int n0 = (ds + 1) * j + i, n1 = n0 + 1, n2 = (ds + 1) * (j + 1) + i, n3 = n2 + 1;
        const int con[2][3] = {{n0, flip ? n1 : n2, flip ? n2 : n1}, {n1, flip ? n3 : n2, flip ? n2 : n3}};
        for (int e = 0; e < 2; ++e) {
          Element *el = new Element(con[e][0], con[e][1], con[e][2]);
          for (int k = 0; k < 3; ++k)
            el->centroid(k) = ((*node[con[e][0]])(k) + (*node[con[e][1]])(k) + (*node[con[e][2]])(k)) / 3.;
          element.Push(el);
        }
      }
    CalcNormals();
    if (flip)
      for (size_t e = 0; e < element.Size(); ++e) element[e]->normal = -element[e]->normal;
    CalcSpheres();
  }

  // unit normal from the winding (Source/Mesh.cpp:229-243)
  void CalcNormals() {
    for (size_t e = 0; e < element.Size(); ++e) {
      const Vec3_t u = *node[element[e]->node[1]] - *node[element[e]->node[0]];
      const Vec3_t v = *node[element[e]->node[2]] - *node[element[e]->node[0]];
      const Vec3_t w = cross(u, v);
      element[e]->normal = w / norm(w);
    }
  }

  // element radius (farthest node from the centroid); the reference's loop
  // leaves nfar at the last node (Source/Mesh.cpp:206-220)
  void CalcSpheres() {
    for (size_t e = 0; e < element.Size(); ++e) {
      double mx = 0.;
      for (int n = 0; n < dimension; ++n) {
        const Vec3_t rv = *node[element[e]->node[n]] - element[e]->centroid;
        if (norm(rv) > mx) mx = norm(rv);
        element[e]->nfar = n;
      }
      element[e]->radius = mx;
    }
    for (size_t e = 0; e < element.Size(); ++e)
      element[e]->pplane = dot(*node[element[e]->node[element[e]->nfar]], element[e]->normal);
  }
};

}  // namespace SPH
