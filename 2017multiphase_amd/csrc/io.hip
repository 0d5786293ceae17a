// Mesh files and the point-cloud reader (host code): outputPovRaywater,
// outputOBJwater, outputBOBJwater (spray.cu:2250-2435) and readdata_solid
// (LoadSolidData.cpp:12-59).  The file layouts are the reference's, so its
// renderers and importers read these files unchanged.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <vector>

#include "mpf_internal.h"

namespace mpf {
namespace {

// verifyfloat3, spray.cu:999-1008: a component that compares false both ways
// with 0 and 1 (NaN) makes the vector invalid; invalid vectors are written as 0
inline void sanitize(float* v) {
  for (int a = 0; a < 3; ++a)
    if (std::isnan(v[a])) {
      v[0] = v[1] = v[2] = 0.0f;
      return;
    }
}

int write_text(const char* path, int fmt, const char* name, std::vector<float>& V, std::vector<float>& N,
               const std::vector<uint32_t>& T, int64_t nv, int64_t nt) {
  FILE* fp = std::fopen(path, "w");
  if (!fp) return MPF_IO;
  if (nv == 0) {  // no mesh: an empty file, as the reference leaves it
    std::fclose(fp);
    return MPF_OK;
  }
  if (fmt == MPF_MESH_POVRAY) {
    std::fprintf(fp, "#declare watermesh=mesh2{\n");
    std::fprintf(fp, "vertex_vectors{\n");
    std::fprintf(fp, "%lld,\n", (long long)nv);
    for (int64_t i = 0; i < nv; ++i) std::fprintf(fp, "< %f, %f, %f >,\n", V[3 * i], V[3 * i + 1], V[3 * i + 2]);
    std::fprintf(fp, "}\n ");
    std::fprintf(fp, "normal_vectors{\n");
    std::fprintf(fp, "%lld,\n", (long long)nv);
    for (int64_t i = 0; i < nv; ++i) std::fprintf(fp, "<%f, %f, %f>,\n", N[3 * i], N[3 * i + 1], N[3 * i + 2]);
    std::fprintf(fp, "}\n ");
    std::fprintf(fp, "face_indices{\n");
    std::fprintf(fp, "%lld,\n", (long long)nt);
    for (int64_t i = 0; i < nt; ++i) std::fprintf(fp, "<%u, %u, %u>,\n", T[3 * i], T[3 * i + 1], T[3 * i + 2]);
    std::fprintf(fp, "}\n ");
    std::fprintf(fp, "inside_vector <0,0,1> }\n ");
    std::fprintf(fp, "object{ watermesh material{%s_material} }\n ", name);
  } else {
    std::fprintf(fp, "#obj model\n");
    std::fprintf(fp, "#vertex_vectors\n");
    for (int64_t i = 0; i < nv; ++i) std::fprintf(fp, "v  %f %f %f\n", V[3 * i], V[3 * i + 1], V[3 * i + 2]);
    std::fprintf(fp, "#normal_vectors\n");
    for (int64_t i = 0; i < nv; ++i) std::fprintf(fp, "vn  %f %f %f\n", N[3 * i], N[3 * i + 1], N[3 * i + 2]);
    std::fprintf(fp, "#face_indices\n");
    for (int64_t i = 0; i < nt; ++i)  // OBJ indices start at 1
      std::fprintf(fp, "f  %d %d %d\n", (int)T[3 * i] + 1, (int)T[3 * i + 1] + 1, (int)T[3 * i + 2] + 1);
  }
  const bool ok = std::ferror(fp) == 0;
  return (std::fclose(fp) == 0 && ok) ? MPF_OK : MPF_IO;
}

// Blender's .bobj.gz: int count + float[3] vertices, int count + float[3]
// normals, int index count + int indices (the reference writes the number
// of indices there, 3 per triangle)
int write_bobj(const char* path, std::vector<float>& V, std::vector<float>& N, const std::vector<uint32_t>& T,
               int64_t nv, int64_t nt) {
  gzFile gz = gzopen(path, "wb1");
  if (!gz) return MPF_IO;
  bool ok = true;
  if (nv > 0) {
    const int n = (int)nv;
    ok = ok && gzwrite(gz, &n, sizeof(int)) == (int)sizeof(int);
    ok = ok && gzwrite(gz, V.data(), (unsigned)(12 * nv)) == (int)(12 * nv);
    ok = ok && gzwrite(gz, &n, sizeof(int)) == (int)sizeof(int);
    ok = ok && gzwrite(gz, N.data(), (unsigned)(12 * nv)) == (int)(12 * nv);
    const int ni = (int)(3 * nt);
    ok = ok && gzwrite(gz, &ni, sizeof(int)) == (int)sizeof(int);
    ok = ok && (nt == 0 || gzwrite(gz, T.data(), (unsigned)(12 * nt)) == (int)(12 * nt));
  }
  return (gzclose(gz) == Z_OK && ok) ? MPF_OK : MPF_IO;
}

}  // namespace

// The last mesh written to path (format MPF_MESH_*); c.err names a failure.
int run_write_mesh(Ctx& c, const char* path, int format, const char* object_name) {
  const int64_t nv = c.mc_nv, nt = c.mc_nt;
  std::vector<float> V((size_t)3 * nv), N((size_t)3 * nv);
  std::vector<uint32_t> T((size_t)3 * nt);
  if (nv) {
    hipMemcpyAsync(V.data(), c.mc_verts, (size_t)nv * 12, hipMemcpyDeviceToHost, c.stream);
    hipMemcpyAsync(N.data(), c.mc_normals, (size_t)nv * 12, hipMemcpyDeviceToHost, c.stream);
  }
  if (nt) hipMemcpyAsync(T.data(), c.mc_tris, (size_t)nt * 12, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "write_mesh")) return MPF_HIP_ERROR;
  for (int64_t i = 0; i < nv; ++i) {
    sanitize(&V[3 * i]);
    sanitize(&N[3 * i]);
  }
  const int st = format == MPF_MESH_BOBJ_GZ
                     ? write_bobj(path, V, N, T, nv, nt)
                     : write_text(path, format, object_name ? object_name : "water", V, N, T, nv, nt);
  if (st != MPF_OK) c.err = std::string("cannot write mesh file ") + path;
  return st;
}
}  // namespace mpf

using namespace mpf;

extern "C" mpf_status mpf_read_point_cloud(const char* path, int64_t cap, double* xyz, int64_t* n) {
  if (!path || !n || cap < 0 || (cap > 0 && !xyz)) {
    g_noctx_err = "null argument";
    return MPF_INVALID_ARG;
  }
  std::ifstream f(path);
  if (!f) {
    g_noctx_err = std::string("cannot open point cloud ") + path;
    return MPF_IO;
  }
  long long count = 0;
  if (!(f >> count) || count < 0) {
    g_noctx_err = std::string("no point count on the first line of ") + path;
    return MPF_IO;
  }
  *n = count;
  // each later line: a point index, then three coordinates (a separator
  // character follows every number, as the reference reads them)
  std::string line;
  std::getline(f, line);
  int64_t i = 0;
  while (i < count && std::getline(f, line)) {
    long long idx;
    double v[3];
    std::istringstream ls(line);
    if (!(ls >> idx)) continue;  // blank line
    ls.get();
    bool ok = true;
    for (int a = 0; a < 3 && ok; ++a) {
      ok = static_cast<bool>(ls >> v[a]);
      ls.get();
    }
    if (!ok) {
      g_noctx_err = std::string("malformed point line in ") + path;
      return MPF_IO;
    }
    if (i < cap)
      for (int a = 0; a < 3; ++a) xyz[3 * i + a] = v[a];
    ++i;
  }
  if (i < count) {
    g_noctx_err = std::string("point cloud ") + path + " ends before its declared count";
    return MPF_IO;
  }
  return MPF_OK;
}
