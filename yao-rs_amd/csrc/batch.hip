// batch.hip — batched VQE energies (many ansatz instances per launch).
#include "kernels.h"
#include "runtime.h"

using namespace yh;

struct yh_vqe_batch {
  int max_instances = 0, max_qubits = 0, dtype = YH_C64;
};

extern "C" {

int yh_vqe_batch_create(int max_instances, int max_qubits, int dtype, yh_vqe_batch** out) {
  return fail(YH_ERR_UNSUPPORTED, "batched VQE is not built yet");
}
int yh_vqe_batch_destroy(yh_vqe_batch* b) {
  delete b;
  return YH_OK;
}
int yh_vqe_batch_energies(yh_vqe_batch* b, int ninst, const int32_t* nq, const int32_t* depth,
                          const int64_t* offset, const double* angles, double J, double h, double* energies) {
  return fail(YH_ERR_UNSUPPORTED, "batched VQE is not built yet");
}
int yh_vqe_plain_energies(int ninst, const int32_t* nq, const int32_t* depth, const int64_t* offset,
                          const double* angles, double J, double h, int dtype, double* energies) {
  return fail(YH_ERR_UNSUPPORTED, "batched VQE is not built yet");
}

}  // extern "C"
