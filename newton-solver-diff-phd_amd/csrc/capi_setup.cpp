// C-ABI of the host-side setup: the exact sparse inverse factor, vertex
// incidence and element geometry (declared in include/diffphd.h).
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/diffphd.h"
#include "error.h"
#include "setup.h"

struct dphd_factor {
  dphd::Factor f;
};

namespace dphd {

thread_local std::string g_last_error;

void set_error(const std::string& message) { g_last_error = message; }

}  // namespace dphd

extern "C" {

const char* dphd_last_error(void) { return dphd::g_last_error.c_str(); }

int dphd_version(void) { return DPHD_ABI_VERSION; }

int dphd_factor_create(int n, const int* offsets, const int* columns, const double* values, dphd_factor** out) {
  if (!out || n < 0 || (n > 0 && (!offsets || !columns || !values))) {
    dphd::set_error("dphd_factor_create: invalid arguments");
    return DPHD_EINVAL;
  }
  try {
    dphd::Csr a;
    a.n = n;
    a.off.assign(offsets, offsets + n + 1);
    a.col.assign(columns, columns + offsets[n]);
    a.val.assign(values, values + offsets[n]);
    auto* h = new dphd_factor;
    h->f = dphd::build_factor(a);
    *out = h;
    return DPHD_OK;
  } catch (const std::exception& e) {
    dphd::set_error(e.what());
    return DPHD_EINVAL;
  }
}

int dphd_factor_sizes(const dphd_factor* h, int64_t* nnz, int64_t* factorization_nnz, int64_t* matrix_nnz) {
  if (!h) {
    dphd::set_error("dphd_factor_sizes: null factor");
    return DPHD_EINVAL;
  }
  if (nnz) *nnz = static_cast<int64_t>(h->f.s.col.size());
  if (factorization_nnz) *factorization_nnz = h->f.factorization_nnz;
  if (matrix_nnz) *matrix_nnz = h->f.matrix_nnz;
  return DPHD_OK;
}

int dphd_factor_export(const dphd_factor* h, int* permutation, int* s_offsets, int* s_columns, double* s_values,
                       int* t_offsets, int* t_columns, double* t_values) {
  if (!h) {
    dphd::set_error("dphd_factor_export: null factor");
    return DPHD_EINVAL;
  }
  const auto& f = h->f;
  const size_t z = f.s.col.size();
  if (permutation) std::memcpy(permutation, f.permutation.data(), sizeof(int) * f.n);
  if (s_offsets) std::memcpy(s_offsets, f.s.off.data(), sizeof(int) * (f.n + 1));
  if (s_columns) std::memcpy(s_columns, f.s.col.data(), sizeof(int) * z);
  if (s_values) std::memcpy(s_values, f.s.val.data(), sizeof(double) * z);
  if (t_offsets) std::memcpy(t_offsets, f.st.off.data(), sizeof(int) * (f.n + 1));
  if (t_columns) std::memcpy(t_columns, f.st.col.data(), sizeof(int) * z);
  if (t_values) std::memcpy(t_values, f.st.val.data(), sizeof(double) * z);
  return DPHD_OK;
}

void dphd_factor_destroy(dphd_factor* h) { delete h; }

int dphd_build_incidence(int vertices, int elements, int corners, const int* indices, int* offsets, int* entries) {
  if (vertices < 0 || elements < 0 || corners <= 0 || !indices || !offsets || !entries) {
    dphd::set_error("dphd_build_incidence: invalid arguments");
    return DPHD_EINVAL;
  }
  const int64_t total = static_cast<int64_t>(elements) * corners;
  for (int64_t t = 0; t < total; ++t)
    if (indices[t] < 0 || indices[t] >= vertices) {
      dphd::set_error("dphd_build_incidence: vertex index out of range");
      return DPHD_EINVAL;
    }
  std::vector<int> off, ent;
  dphd::build_incidence(vertices, elements, corners, indices, off, ent);
  std::memcpy(offsets, off.data(), sizeof(int) * (vertices + 1));
  std::memcpy(entries, ent.data(), sizeof(int) * total);
  return DPHD_OK;
}

int dphd_tet_geometry(int elements, const float* poses, double* gradients, double* volumes) {
  if (elements < 0 || !poses || !gradients || !volumes) {
    dphd::set_error("dphd_tet_geometry: invalid arguments");
    return DPHD_EINVAL;
  }
  dphd::tet_geometry(elements, poses, gradients, volumes);
  return DPHD_OK;
}

int dphd_hex_gradients(double dx, double* out) {
  if (!(dx > 0.0) || !out) {
    dphd::set_error("dphd_hex_gradients: invalid arguments");
    return DPHD_EINVAL;
  }
  dphd::hex_gradients(dx, out);
  return DPHD_OK;
}

int dphd_assemble(int vertices, int elements, int corners, int samples, const int* indices, const double* grads,
                  int shared_grads, const double* weights, const double* fibre_dir, const double* fibre_w,
                  const unsigned char* free, double inertia, int plane_count, const int* plane_vertices,
                  const double* plane_stiffness, int* nnz, int* offsets, int* columns, double* values, int capacity) {
  try {
    dphd::AssemblyInput in;
    in.vertices = vertices;
    in.elements = elements;
    in.corners = corners;
    in.samples = samples;
    in.indices = indices;
    in.grads = grads;
    in.shared_grads = shared_grads != 0;
    in.weights = weights;
    in.fibre_dir = fibre_dir;
    in.fibre_w = fibre_w;
    in.free = free;
    in.inertia = inertia;
    in.plane_count = plane_count;
    in.plane_vertices = plane_vertices;
    in.plane_stiffness = plane_stiffness;
    const dphd::Csr a = dphd::assemble_matrix(in);
    *nnz = static_cast<int>(a.col.size());
    if (capacity >= *nnz && offsets && columns && values) {
      std::memcpy(offsets, a.off.data(), sizeof(int) * (vertices + 1));
      std::memcpy(columns, a.col.data(), sizeof(int) * a.col.size());
      std::memcpy(values, a.val.data(), sizeof(double) * a.val.size());
    }
    return DPHD_OK;
  } catch (const std::exception& e) {
    dphd::set_error(e.what());
    return DPHD_EINVAL;
  }
}

}  // extern "C"
