// scene_io.cpp — scene files into the host scene builder: the .xenon scene
// format (scene/scene_file.cpp:49-240), .mat material files
// (material/material.cpp:31-152) and Wavefront OBJ meshes
// (geometry/mesh.cpp:42-151), with the same directives, defaults, path rules
// and arithmetic, so that a file loads into the scene the reference builds
// from it, analytic primitives (sphere, box, disk, quad, plane) included.
//
// Compiled with g++ -ffp-contract=off (the reference's host arithmetic).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "scene.h"
#include "xenon_hip.h"

namespace {

namespace fs = std::filesystem;

std::string strip_quotes(const std::string& s) {
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
  return s;
}

// A relative path as given if it exists, else against the scene file's
// directory (scene_file.cpp:20-27).
std::string beside(const std::string& scene_path, const std::string& rel) {
  if (fs::exists(rel)) return rel;
  const fs::path p = fs::path(scene_path).parent_path() / rel;
  if (fs::exists(p)) return p.string();
  return rel;
}

// Material authoring parameters in the 16-float order of the builder, with
// the defaults of material.h:21-37.
struct MatParams {
  std::string name = "default";
  float p[16] = {0.5f, 0.5f, 0.5f, 0.5f, 0.0f, 1.5f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
};

// .mat file: `key value...` lines, '#' comments (material.cpp:31-104).
// Keys the shading model does not use (subsurfaceMFP, clearcoat,
// clearcoatRoughness, emissionTemperature, type) are read and ignored.
bool read_mat(const std::string& path, MatParams& m) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  std::string line;
  while (std::getline(f, line)) {
    const size_t first = line.find_first_not_of(" \t");
    if (first == std::string::npos || line[first] == '#') continue;
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    float ignore[3];
    if (key == "name") {
      std::string v;
      ss >> v;
      m.name = strip_quotes(v);
    } else if (key == "baseColor") {
      ss >> m.p[0] >> m.p[1] >> m.p[2];
    } else if (key == "roughness") {
      ss >> m.p[3];
    } else if (key == "metallic") {
      ss >> m.p[4];
    } else if (key == "ior") {
      ss >> m.p[5];
    } else if (key == "transmission") {
      ss >> m.p[6];
    } else if (key == "subsurface") {
      ss >> m.p[7];
    } else if (key == "subsurfaceColor") {
      ss >> m.p[8] >> m.p[9] >> m.p[10];
    } else if (key == "anisotropy") {
      ss >> m.p[11];
    } else if (key == "emission") {
      ss >> m.p[12] >> m.p[13] >> m.p[14];
    } else if (key == "subsurfaceMFP") {
      ss >> ignore[0] >> ignore[1] >> ignore[2];
    }
  }
  return true;
}

// OBJ: v, vn (normalised), usemtl, f with pos[/uv[/normal]] corners fanned
// into triangles; one mesh vertex per distinct (position, normal) pair in
// first-use order; corners without a normal get (0,0,0); a file without any
// vn gets area-weighted smooth normals before the mesh is transformed
// (mesh.cpp:42-151).
bool read_obj(const std::string& path, const std::map<std::string, int>& mats, xnh::HostMesh& m) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  std::vector<float> P, Nn;
  std::unordered_map<int64_t, int32_t> seen;
  int cur = 0;
  bool has_normals = false;
  std::string line;
  auto vertex = [&](int pi, int ni) -> int32_t {
    const int64_t key = (static_cast<int64_t>(pi) << 32) | static_cast<uint32_t>(ni);
    auto it = seen.find(key);
    if (it != seen.end()) return it->second;
    const int32_t id = static_cast<int32_t>(m.vx.size());
    m.vx.push_back(P[3 * static_cast<size_t>(pi)]);
    m.vy.push_back(P[3 * static_cast<size_t>(pi) + 1]);
    m.vz.push_back(P[3 * static_cast<size_t>(pi) + 2]);
    const bool hn = ni >= 0 && 3 * static_cast<size_t>(ni) + 2 < Nn.size();
    m.nx.push_back(hn ? Nn[3 * static_cast<size_t>(ni)] : 0.f);
    m.ny.push_back(hn ? Nn[3 * static_cast<size_t>(ni) + 1] : 0.f);
    m.nz.push_back(hn ? Nn[3 * static_cast<size_t>(ni) + 2] : 0.f);
    seen[key] = id;
    return id;
  };
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    char type[8] = {};
    if (std::sscanf(line.c_str(), "%7s", type) < 1) continue;
    const char* p = line.c_str() + std::strlen(type);
    if (std::strcmp(type, "v") == 0) {
      float x, y, z;
      if (std::sscanf(p, " %f %f %f", &x, &y, &z) == 3) {
        P.push_back(x);
        P.push_back(y);
        P.push_back(z);
      }
    } else if (std::strcmp(type, "vn") == 0) {
      float x, y, z;
      if (std::sscanf(p, " %f %f %f", &x, &y, &z) == 3) {
        const float r = 1.f / std::sqrt(x * x + y * y + z * z);   // Vec3 / s multiplies by 1/s
        Nn.push_back(x * r);
        Nn.push_back(y * r);
        Nn.push_back(z * r);
        has_normals = true;
      }
    } else if (std::strcmp(type, "usemtl") == 0) {
      char name[128];
      if (std::sscanf(p, " %127s", name) == 1) {
        auto it = mats.find(name);
        if (it != mats.end()) cur = it->second;
      }
    } else if (std::strcmp(type, "f") == 0) {
      std::vector<int32_t> corners;
      const char* q = p;
      while (*q) {
        while (*q == ' ' || *q == '\t') ++q;
        if (!*q || *q == '\n' || *q == '\r') break;
        int pi = -1, ti = -1, ni = -1, used = 0;
        if (std::sscanf(q, "%d/%d/%d%n", &pi, &ti, &ni, &used) >= 3) {
        } else if (std::sscanf(q, "%d//%d%n", &pi, &ni, &used) >= 2) {
        } else if (std::sscanf(q, "%d/%d%n", &pi, &ti, &used) >= 2) {
        } else if (std::sscanf(q, "%d%n", &pi, &used) >= 1) {
        } else {
          break;
        }
        if (pi > 0) {
          --pi;
          ni = ni > 0 ? ni - 1 : -1;
          // a corner naming a position not yet read: the reference reads out
          // of bounds here; this loader refuses the file
          if (3 * static_cast<size_t>(pi) + 2 >= P.size()) return false;
          corners.push_back(vertex(pi, ni));
        }
        q += used;
      }
      for (size_t i = 1; i + 1 < corners.size(); ++i) {
        m.indices.push_back(corners[0]);
        m.indices.push_back(corners[i]);
        m.indices.push_back(corners[i + 1]);
        m.mat_ids.push_back(cur);
      }
    }
  }
  if (!has_normals) xnh::smooth_normals(m);
  return true;
}

// TriangleMesh::transform (mesh.cpp:156-205): scale, rotate about Y, then X,
// then Z (degrees), translate; normals rotated and renormalised.
void transform(xnh::HostMesh& m, const float t[3], float scale, const float rot_deg[3]) {
  const float k = static_cast<float>(M_PI / 180.f);
  const float rx = rot_deg[0] * k, ry = rot_deg[1] * k, rz = rot_deg[2] * k;
  const float cx = std::cos(rx), sx = std::sin(rx), cy = std::cos(ry), sy = std::sin(ry), cz = std::cos(rz),
              sz = std::sin(rz);
  for (size_t i = 0; i < m.vx.size(); ++i) {
    float x = m.vx[i] * scale, y = m.vy[i] * scale, z = m.vz[i] * scale;
    float nx = x * cy + z * sy, nz = -x * sy + z * cy;
    x = nx;
    z = nz;
    float ny = y * cx - z * sx;
    nz = y * sx + z * cx;
    y = ny;
    z = nz;
    nx = x * cz - y * sz;
    ny = x * sz + y * cz;
    m.vx[i] = nx + t[0];
    m.vy[i] = ny + t[1];
    m.vz[i] = z + t[2];
  }
  for (size_t i = 0; i < m.nx.size(); ++i) {
    float x = m.nx[i], y = m.ny[i], z = m.nz[i];
    float tx = x * cy + z * sy, tz = -x * sy + z * cy;
    x = tx;
    z = tz;
    float ty = y * cx - z * sx;
    tz = y * sx + z * cx;
    y = ty;
    z = tz;
    tx = x * cz - y * sz;
    ty = x * sz + y * cz;
    const float r = 1.f / std::sqrt(tx * tx + ty * ty + z * z);
    m.nx[i] = tx * r;
    m.ny[i] = ty * r;
    m.nz[i] = z * r;
  }
}

int add_host_mesh(xnh_scene* s, const xnh::HostMesh& m) {
  const size_t nv = m.vx.size();
  std::vector<float> pos(3 * nv), nrm(3 * nv);
  for (size_t i = 0; i < nv; ++i) {
    pos[3 * i] = m.vx[i];
    pos[3 * i + 1] = m.vy[i];
    pos[3 * i + 2] = m.vz[i];
    nrm[3 * i] = m.nx[i];
    nrm[3 * i + 1] = m.ny[i];
    nrm[3 * i + 2] = m.nz[i];
  }
  return xnh_scene_add_mesh(s, static_cast<int32_t>(nv), pos.data(), nrm.data(), m.tris(), m.indices.data(),
                            m.mat_ids.data());
}

// n / |n| when |n| > 1e-6 (the parser's normalisation of plane and disk
// normals; Vec3 / s multiplies by 1 / s).
void unit_if_long(float* n) {
  const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (len > 1e-6f) {
    const float r = 1.f / len;
    n[0] *= r;
    n[1] *= r;
    n[2] *= r;
  }
}

}  // namespace

extern "C" {

int xnh_material_load(const char* path, float* params16, char* name, int32_t name_cap) {
  if (!path || !params16) return XNH_E_INVALID;
  MatParams m;
  const bool ok = read_mat(path, m);
  std::memcpy(params16, m.p, sizeof m.p);
  if (name && name_cap > 0) {
    std::strncpy(name, m.name.c_str(), static_cast<size_t>(name_cap) - 1);
    name[name_cap - 1] = '\0';
  }
  return ok ? XNH_OK : XNH_E_INVALID;
}

int xnh_scene_load_file(const char* path, xnh_scene** out, xnh_camera* camera, xnh_file_info* info) {
  if (!path || !out) return XNH_E_INVALID;
  *out = nullptr;
  std::ifstream f(path);
  if (!f.is_open()) return XNH_E_INVALID;
  xnh_scene* s = xnh_scene_new();
  if (!s) return XNH_E_ALLOC;
  xnh_file_info fi{};
  fi.width = 800;
  fi.height = 600;
  fi.samples = 64;
  fi.min_bounces = 3;
  fi.max_bounces = 8;
  // Environment defaults (scene/environment.h:27-31)
  xnh_sky sky = {{0.10f, 0.30f, 0.80f}, {0.55f, 0.70f, 0.85f}, {0.10f, 0.08f, 0.05f}, 3.0f, 1.0f};
  float eye[3] = {0.f, 5.f, 20.f}, target[3] = {0.f, 5.f, 0.f}, up[3] = {0.f, 1.f, 0.f}, fov = 40.f;
  std::map<std::string, int> mats;
  int nmat = 0, nmesh = 0, ninst = 0;
  std::string line;
  int rc = XNH_OK;
  auto add_material = [&](const MatParams& m) {
    // a later material of the same name takes the name (scene_file.cpp:87-88)
    mats[m.name] = nmat++;
    return xnh_scene_add_material(s, m.p);
  };
  while (rc >= 0 && std::getline(f, line)) {
    std::stringstream ss(line);
    std::string cmd;
    ss >> cmd;
    if (cmd.empty() || cmd[0] == '#') continue;
    if (cmd == "config") {
      ss >> fi.width >> fi.height >> fi.samples;
    } else if (cmd == "camera") {
      ss >> eye[0] >> eye[1] >> eye[2] >> target[0] >> target[1] >> target[2] >> fov;
    } else if (cmd == "matfile") {
      std::string m;
      ss >> m;
      MatParams mp;
      if (!read_mat(beside(path, "materials/" + strip_quotes(m)), mp)) ++fi.missing_files;
      rc = add_material(mp);
    } else if (cmd == "material") {
      // inline form: name r g b metallic roughness [specular ior transmission
      // transmission_roughness] (scene_file.cpp:92-104, material.cpp:134-152)
      MatParams mp;
      float metallic = 0.f, rough = 0.f, spec = 0.5f, ior = 1.5f, trans = 0.f, trough = -1.f;
      ss >> mp.name >> mp.p[0] >> mp.p[1] >> mp.p[2] >> metallic >> rough;
      ss >> spec >> ior >> trans >> trough;
      mp.p[3] = (trough >= 0.f && trans > 0.01f) ? trough : rough;
      mp.p[4] = metallic;
      mp.p[5] = ior;
      mp.p[6] = trans;
      rc = add_material(mp);
    } else if (cmd == "mesh") {
      // opened as given (relative to the working directory), unlike
      // matfile paths (scene_file.cpp:107-113)
      std::string obj;
      ss >> obj;
      xnh::HostMesh m;
      if (!read_obj(obj, mats, m)) {
        ++fi.missing_files;
        continue;
      }
      float t[3] = {0.f, 0.f, 0.f}, rot[3] = {0.f, 0.f, 0.f}, scale = 1.f;
      if (ss >> t[0] >> t[1] >> t[2] >> scale >> rot[0] >> rot[1] >> rot[2]) transform(m, t, scale, rot);
      rc = add_host_mesh(s, m);
      if (rc >= 0) ++nmesh;
    } else if (cmd == "light") {
      int32_t mesh = 0, tri = 0;
      float e[3] = {0.f, 0.f, 0.f}, area = 0.f;
      ss >> mesh >> tri >> e[0] >> e[1] >> e[2] >> area;
      rc = xnh_scene_add_light(s, mesh, tri, e, area);
    } else if (cmd == "sunlight") {
      float d[3] = {0.f, 0.f, 0.f}, c[3] = {0.f, 0.f, 0.f}, intensity = 0.f;
      ss >> d[0] >> d[1] >> d[2] >> c[0] >> c[1] >> c[2] >> intensity;
      rc = xnh_scene_add_dir_light(s, d, c, intensity);
    } else if (cmd == "sky") {
      // zenith and horizon always; ground, then intensity and sharpness
      // (1 and 3 unless given), only when a ground colour follows
      // (scene_file.cpp:154-172)
      float z[3] = {0.f, 0.f, 0.f}, h[3] = {0.f, 0.f, 0.f};
      ss >> z[0] >> z[1] >> z[2] >> h[0] >> h[1] >> h[2];
      std::memcpy(sky.zenith, z, sizeof z);
      std::memcpy(sky.horizon, h, sizeof h);
      float g[3] = {0.1f, 0.08f, 0.05f}, intensity = 1.f, sharp = 3.f;
      if (ss >> g[0] >> g[1] >> g[2]) {
        std::memcpy(sky.ground, g, sizeof g);
        ss >> intensity >> sharp;
        sky.intensity = intensity;
        sky.sharpness = sharp;
      }
      rc = xnh_scene_set_sky(s, sky.zenith, sky.horizon, sky.ground, sky.intensity, sky.sharpness);
    } else if (cmd == "instance") {
      // instance <mesh> <tx> <ty> <tz> <scale> <rx> <ry> <rz> [id]: a TRS
      // placement of a loaded mesh (no counterpart in the reference's
      // format, which bakes transforms into vertices; its parser skips the
      // line). With any instance line the TLAS holds exactly these.
      int32_t mesh = -1;
      float t[3] = {0.f, 0.f, 0.f}, rot[3] = {0.f, 0.f, 0.f}, scale = 1.f;
      ss >> mesh >> t[0] >> t[1] >> t[2] >> scale >> rot[0] >> rot[1] >> rot[2];
      if (!ss || mesh < 0 || mesh >= nmesh) {
        rc = XNH_E_INVALID;
        break;
      }
      int32_t id = ninst;
      if (!(ss >> id)) id = ninst;
      rc = xnh_scene_add_instance(s, mesh, t, scale, rot, id);
      ++ninst;
    } else if (cmd == "sphere" || cmd == "box" || cmd == "disk" || cmd == "quad" || cmd == "plane") {
      // analytic primitives (scene_file.cpp:175-233); an unknown material
      // name selects material 0, as resolve_mat does
      float q[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      std::string mname;
      int type = XNH_PRIM_SPHERE;
      if (cmd == "sphere") {
        ss >> q[0] >> q[1] >> q[2] >> q[3] >> mname;
      } else if (cmd == "box") {
        type = XNH_PRIM_BOX;
        ss >> q[0] >> q[1] >> q[2] >> q[3] >> q[4] >> q[5] >> mname;
      } else if (cmd == "disk") {
        type = XNH_PRIM_DISK;
        ss >> q[0] >> q[1] >> q[2] >> q[3] >> q[4] >> q[5] >> q[6] >> mname;
        unit_if_long(q + 3);
      } else if (cmd == "quad") {
        type = XNH_PRIM_QUAD;
        ss >> q[0] >> q[1] >> q[2] >> q[3] >> q[4] >> q[5] >> q[6] >> q[7] >> q[8] >> mname;
      } else {
        type = XNH_PRIM_PLANE;
        ss >> q[0] >> q[1] >> q[2] >> q[3] >> mname;
        unit_if_long(q);
      }
      auto it = mats.find(mname);
      int mid = 0;
      if (it != mats.end()) mid = it->second;
      else ++fi.unknown_materials;
      rc = xnh_scene_add_primitive(s, type, q, mid);
      if (rc >= 0) ++fi.primitives;
    }
  }
  if (rc >= 0) rc = xnh_scene_build(s, 0);
  if (rc < 0) {
    xnh_scene_free(s);
    return rc;
  }
  fi.meshes = nmesh;
  fi.instances = ninst;
  std::memcpy(fi.eye, eye, sizeof eye);
  std::memcpy(fi.target, target, sizeof target);
  std::memcpy(fi.up, up, sizeof up);
  fi.fov = fov;
  if (camera) {
    const float aspect = static_cast<float>(fi.width) / static_cast<float>(fi.height);
    xnh_camera_look_at(eye, target, up, fov, aspect, camera);
  }
  if (info) *info = fi;
  *out = s;
  return XNH_OK;
}

}  // extern "C"
