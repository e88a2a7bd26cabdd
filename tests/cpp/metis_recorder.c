/* A stand-in libmetis for tests: exports METIS_PartGraphKway with METIS's
 * default 32-bit idx_t / real_t signature, writes every argument it
 * receives to the text file named by FAP_METIS_RECORD, and returns the
 * partition part[i] = (7 i) mod nparts with METIS_OK (1). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int METIS_PartGraphKway(int32_t *nvtxs, int32_t *ncon, int32_t *xadj, int32_t *adjncy,
                        int32_t *vwgt, int32_t *vsize, int32_t *adjwgt, int32_t *nparts,
                        float *tpwgts, float *ubvec, int32_t *options, int32_t *objval,
                        int32_t *part) {
  const char *path = getenv("FAP_METIS_RECORD");
  FILE *f = path ? fopen(path, "w") : NULL;
  if (!f) return -4; /* METIS_ERROR */
  fprintf(f, "%d %d %d\n", *nvtxs, *ncon, *nparts);
  fprintf(f, "%d %d %d %d %d %d\n", vwgt == NULL, vsize == NULL, adjwgt == NULL, tpwgts == NULL,
          ubvec == NULL, options == NULL);
  for (int32_t i = 0; i <= *nvtxs; i++) fprintf(f, "%d ", xadj[i]);
  fprintf(f, "\n");
  for (int32_t i = 0; i < xadj[*nvtxs]; i++) fprintf(f, "%d ", adjncy[i]);
  fprintf(f, "\n");
  fclose(f);
  for (int32_t i = 0; i < *nvtxs; i++) part[i] = (int32_t)((7 * (int64_t)i) % *nparts);
  *objval = 0;
  return 1;
}
