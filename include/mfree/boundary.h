// src/boundary.h of the reference: the outline of a body (particle indexes in
// order around the body, their positions and outward normals), which the
// penalty contact of src/contact.h uses.  Positions are those of the listed
// particles; normals are the mean of the two adjacent edge normals.
#ifndef MFREE_BOUNDARY_H_
#define MFREE_BOUNDARY_H_

#include <math.h>

#include <vector>

template <typename T>
class boundary {
  std::vector<unsigned int> m_idx;
  std::vector<double> m_x, m_y, m_nx, m_ny;

 public:
  boundary() {}
  boundary(std::vector<unsigned int> indexes, T** particles) : m_idx(indexes) {
    update(particles);
  }

  // positions of the boundary particles and their normals
  void update(T** particles) {
    const size_t n = m_idx.size();
    m_x.resize(n);
    m_y.resize(n);
    for (size_t k = 0; k < n; k++) {
      m_x[k] = particles[m_idx[k]]->x;
      m_y[k] = particles[m_idx[k]]->y;
    }
    update_normals();
  }

  void update_normals() {
    const size_t n = m_idx.size();
    m_nx.assign(n, 0.);
    m_ny.assign(n, 0.);
    for (size_t k = 0; k < n && n > 1; k++) {
      const size_t l = (k + n - 1) % n, r = (k + 1) % n;
      // counter-clockwise outline: the outward normal of the edge from the
      // left neighbour is (-tl.y, tl.x) with tl = left - p, that of the edge to
      // the right neighbour (tr.y, -tr.x) with tr = right - p; each scaled by
      // 1/sqrt(|v|^2) (glm::normalize, as the GPU does)
      const double tlx = m_x[l] - m_x[k], tly = m_y[l] - m_y[k];
      const double trx = m_x[r] - m_x[k], try_ = m_y[r] - m_y[k];
      const double ax = -tly, ay = tlx, bx = try_, by = -trx;
      const double ia = 1. / sqrt(ax * ax + ay * ay), ib = 1. / sqrt(bx * bx + by * by);
      m_nx[k] = 0.5 * (ax * ia + bx * ib);
      m_ny[k] = 0.5 * (ay * ia + by * ib);
    }
  }

  unsigned int num() const { return (unsigned int)m_idx.size(); }
  const std::vector<unsigned int>& indexes() const { return m_idx; }
  const std::vector<double>& x() const { return m_x; }
  const std::vector<double>& y() const { return m_y; }
  const std::vector<double>& nx() const { return m_nx; }
  const std::vector<double>& ny() const { return m_ny; }
};

#endif
