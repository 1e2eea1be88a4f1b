// Launch arguments of the contact kernels (device pointers; the tool by value).
#pragma once

#include <hip/hip_runtime.h>

namespace sbk {

struct ContactTool {
  int shape;  // 0 plane, 1 sphere, 2 cylinder
  double centre[3], centre0[3];
  double radius;
  double normal[3], axis[3];
  double friction, tangential_ratio;
};

struct ContactArgs {
  ContactTool tool;
  int friction_symmetric;
  int want_tangent;
  const double* coords;        // 3 per node
  const double* u;             // 3 per node
  const double* start;         // u at the start of the increment, 3 per node
  long long num_slaves;
  const int* nodes;            // per slave
  const double* kappa_area;    // kappa_j A_j per slave
  const int* hist_present;     // committed friction history per slave (may be null: none)
  const double* hist_force;
  const double* hist_relative;
  double* residual;            // 3 per node, decremented by the contact forces
  double* blocks;              // 9 per slave (row-major)
  int* active;                 // per slave
  int* trial_present;          // trial friction history per slave
  double* trial_force;
  double* trial_relative;
  double* node_terms;          // 8 per slave: terms of the norms and resultants
  int* status;                 // per slave: 1 = through the tool (SolverError)
};

int launch_contact(const ContactArgs& a, double* scalars, hipStream_t s);
int launch_add_node_blocks(const double* blocks, const int* active, const int* first_value, const int* row_len,
                           long long ns, int column_major, double* values, hipStream_t s);
// work: ns + ceil(ns / 256) int32
int launch_contact_candidates(const ContactTool& t, const double* coords, const double* u0, long long ns,
                              const int* nodes, double margin, int* candidates, int* count, int* work,
                              hipStream_t s);
int launch_contact_anticipate(const ContactTool& t, const double* coords, const double* u, const double* du,
                              long long ns, const int* nodes, const double* kappa_area, double* residual,
                              double* blocks, int* active, int* added, hipStream_t s);

}  // namespace sbk
