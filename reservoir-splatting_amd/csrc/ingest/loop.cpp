// Loop subdivision of a triangle mesh (Loop 1987; the scheme of pbrt's "loopsubdiv"
// shape, which the reference imports through Source/plugins/importers/PBRTImporter/
// LoopSubdivide.cpp): each level splits every triangle into four, moves each existing
// ("even") vertex to (1 - n beta) p + beta sum(ring) with beta = 3/16 for valence 3 and
// 3/(8n) otherwise (boundary vertices: 3/4 p + 1/8 of the two boundary neighbours), and
// places each new ("odd") edge vertex at 3/8 of the edge ends plus 1/8 of the two
// opposite vertices (boundary edges: the midpoint).  After the last level the vertices
// are pushed to the limit surface (gamma = 1/(n + 3/(8 beta)); boundary: 3/5 p + 1/5 of
// the boundary neighbours) and the limit normals come from the two tangents of the
// one-ring (cos/sin-weighted sums; the boundary tangent mask across the ring).
#include <algorithm>
#include <array>
#include <unordered_map>

#include "ingest.h"

namespace ingest {

namespace {

constexpr double kPiD = 3.14159265358979323846;

using Tri = std::array<uint32_t, 3>;

uint64_t edge_key(uint32_t a, uint32_t b) { return ((uint64_t)std::min(a, b) << 32) | std::max(a, b); }

struct Topology {
  // For each directed edge (a -> b) of a face: the face and the local index of a.
  std::unordered_map<uint64_t, std::pair<uint32_t, int>> directed;
  std::vector<uint32_t> vertexFace;  // one face containing each vertex
  std::vector<int> vertexLocal;

  Topology(size_t nv, const std::vector<Tri>& F) : vertexFace(nv, UINT32_MAX), vertexLocal(nv, -1) {
    for (uint32_t f = 0; f < F.size(); f++)
      for (int i = 0; i < 3; i++) {
        uint64_t k = ((uint64_t)F[f][i] << 32) | F[f][(i + 1) % 3];
        if (!directed.emplace(k, std::make_pair(f, i)).second)
          throw std::runtime_error("loopsubdiv: non-manifold or inconsistently wound mesh");
        if (vertexFace[F[f][i]] == UINT32_MAX) {
          vertexFace[F[f][i]] = f;
          vertexLocal[F[f][i]] = i;
        }
      }
  }

  // The face across the directed edge a -> b (the one holding b -> a), or UINT32_MAX.
  bool opposite(uint32_t a, uint32_t b, uint32_t& face, int& local) const {
    auto it = directed.find(((uint64_t)b << 32) | a);
    if (it == directed.end()) return false;
    face = it->second.first;
    local = it->second.second;
    return true;
  }

  // Ordered one-ring of v: for an interior vertex the ring around it; for a boundary
  // vertex the ring from one boundary neighbour to the other.
  std::vector<uint32_t> ring(uint32_t v, const std::vector<Tri>& F, bool& boundary) const {
    boundary = false;
    uint32_t f0 = vertexFace[v];
    if (f0 == UINT32_MAX) return {};
    int i0 = vertexLocal[v];
    // Walk backwards to a boundary face if there is one.
    uint32_t f = f0;
    int i = i0;
    for (size_t guard = 0; guard < 4096; guard++) {
      uint32_t vn = F[f][(i + 1) % 3];
      uint32_t g;
      int gl;
      // The face before f around v holds the edge vn -> v.
      if (!opposite(v, vn, g, gl)) {
        boundary = true;
        break;
      }
      int gi = (gl + 1) % 3;  // local index of v in g
      f = g;
      i = gi;
      if (f == f0) break;
    }
    std::vector<uint32_t> r;
    uint32_t start = f;
    r.push_back(F[f][(i + 1) % 3]);
    for (size_t guard = 0; guard < 4096; guard++) {
      uint32_t vp = F[f][(i + 2) % 3];
      uint32_t g;
      int gl;
      if (!opposite(vp, v, g, gl)) {  // no face across v -> vp's twin: the ring ends here
        r.push_back(vp);
        break;
      }
      f = g;
      i = gl;  // in g the edge is v -> vp, so v is at gl
      if (f == start) break;
      r.push_back(F[f][(i + 1) % 3]);
    }
    return r;
  }
};

double beta(size_t n) { return n == 3 ? 3.0 / 16.0 : 3.0 / (8.0 * (double)n); }

f3 weighted(const std::vector<f3>& P, const std::vector<uint32_t>& ring, double wSelf, f3 self, double wRing) {
  double x = wSelf * self.x, y = wSelf * self.y, z = wSelf * self.z;
  for (uint32_t r : ring) {
    x += wRing * P[r].x;
    y += wRing * P[r].y;
    z += wRing * P[r].z;
  }
  return {(float)x, (float)y, (float)z};
}

}  // namespace

Mesh loop_subdivide_mesh(int levels, const std::vector<f3>& P0, const std::vector<uint32_t>& indices) {
  if (indices.size() % 3) throw std::runtime_error("loopsubdiv: index count not a multiple of 3");
  for (uint32_t i : indices)
    if (i >= P0.size()) throw std::runtime_error("loopsubdiv: vertex index out of range");
  std::vector<f3> P = P0;
  std::vector<Tri> F;
  for (size_t t = 0; t < indices.size(); t += 3) F.push_back({indices[t], indices[t + 1], indices[t + 2]});

  for (int level = 0; level < levels; level++) {
    Topology topo(P.size(), F);
    std::vector<f3> next(P.size());
    // Even vertices.
    for (uint32_t v = 0; v < P.size(); v++) {
      bool boundary;
      std::vector<uint32_t> r = topo.ring(v, F, boundary);
      if (r.empty()) {
        next[v] = P[v];
      } else if (boundary) {
        next[v] = weighted(P, {r.front(), r.back()}, 0.75, P[v], 0.125);
      } else {
        double b = beta(r.size());
        next[v] = weighted(P, r, 1.0 - (double)r.size() * b, P[v], b);
      }
    }
    // Odd vertices, one per edge.
    std::unordered_map<uint64_t, uint32_t> edgeVertex;
    for (uint32_t f = 0; f < F.size(); f++)
      for (int i = 0; i < 3; i++) {
        uint32_t a = F[f][i], b = F[f][(i + 1) % 3];
        uint64_t k = edge_key(a, b);
        if (edgeVertex.count(k)) continue;
        uint32_t c = F[f][(i + 2) % 3];
        uint32_t g;
        int gl;
        f3 p;
        if (topo.opposite(a, b, g, gl)) {
          uint32_t d = F[g][(gl + 2) % 3];
          p = weighted(P, {a, b}, 0.0, P[a], 0.375);
          p = p + (P[c] + P[d]) * 0.125f;
        } else {
          p = (P[a] + P[b]) * 0.5f;
        }
        edgeVertex.emplace(k, (uint32_t)next.size());
        next.push_back(p);
      }
    std::vector<Tri> nf;
    nf.reserve(F.size() * 4);
    for (const Tri& t : F) {
      uint32_t e01 = edgeVertex[edge_key(t[0], t[1])], e12 = edgeVertex[edge_key(t[1], t[2])],
               e20 = edgeVertex[edge_key(t[2], t[0])];
      nf.push_back({t[0], e01, e20});
      nf.push_back({t[1], e12, e01});
      nf.push_back({t[2], e20, e12});
      nf.push_back({e01, e12, e20});
    }
    P = std::move(next);
    F = std::move(nf);
  }

  // Limit positions and normals.
  Topology topo(P.size(), F);
  Mesh out;
  out.name = "loopsubdiv";
  out.positions.resize(P.size());
  out.normals.resize(P.size());
  std::vector<f3> faceAcc(P.size());
  for (const Tri& t : F) {
    f3 n = cross(P[t[1]] - P[t[0]], P[t[2]] - P[t[0]]);
    for (uint32_t v : t) faceAcc[v] = faceAcc[v] + n;
  }
  for (uint32_t v = 0; v < P.size(); v++) {
    bool boundary;
    std::vector<uint32_t> r = topo.ring(v, F, boundary);
    size_t n = r.size();
    if (n == 0) {
      out.positions[v] = P[v];
      out.normals[v] = normalize(faceAcc[v]);
      continue;
    }
    f3 S, T;
    if (boundary) {
      out.positions[v] = weighted(P, {r.front(), r.back()}, 0.6, P[v], 0.2);
      S = P[r.back()] - P[r.front()];
      if (n == 2) {
        T = P[r[0]] + P[r[1]] - P[v] * 2.f;
      } else if (n == 3) {
        T = P[r[1]] - P[v];
      } else if (n == 4) {
        T = P[r[0]] * -1.f + P[r[1]] * 2.f + P[r[2]] * 2.f + P[r[3]] * -1.f + P[v] * -2.f;
      } else {
        double theta = kPiD / (double)(n - 1);
        f3 acc = (P[r[0]] + P[r[n - 1]]) * (float)std::sin(theta);
        for (size_t k = 1; k + 1 < n; k++) {
          double wt = (2.0 * std::cos(theta) - 2.0) * std::sin((double)k * theta);
          acc = acc + P[r[k]] * (float)wt;
        }
        T = acc * -1.f;
      }
    } else {
      double b = beta(n);
      double gamma = 1.0 / ((double)n + 3.0 / (8.0 * b));
      out.positions[v] = weighted(P, r, 1.0 - (double)n * gamma, P[v], gamma);
      for (size_t k = 0; k < n; k++) {
        S = S + P[r[k]] * (float)std::cos(2.0 * kPiD * (double)k / (double)n);
        T = T + P[r[k]] * (float)std::sin(2.0 * kPiD * (double)k / (double)n);
      }
    }
    f3 N = normalize(cross(S, T));
    if (dot(N, faceAcc[v]) < 0.f) N = N * -1.f;  // orient with the winding
    if (length(N) == 0.f) N = normalize(faceAcc[v]);
    out.normals[v] = N;
  }
  for (const Tri& t : F) out.indices.insert(out.indices.end(), t.begin(), t.end());
  return out;
}

}  // namespace ingest
