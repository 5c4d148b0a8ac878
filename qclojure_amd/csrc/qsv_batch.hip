// qsv_batch.hip — many small circuits in one persistent launch (placeholder
// until the batched engine lands: reports the call as unsupported).
#include "qsv_internal.h"

extern "C" int qsv_batch_expect_z(const qsv_batch_desc* desc, double* energies)
{
  (void)desc;
  (void)energies;
  return QSV_E_UNSUPPORTED;
}
