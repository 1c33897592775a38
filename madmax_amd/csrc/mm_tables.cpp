// mm_tables.cpp -- host construction of the structured hex8 slot tables.
//
// The reference fills its ELL by calling ell_add_vals once per element in
// ascending element id (micro/src/assembly.c:93-110); ell_add_val puts a new
// column in the row's first free slot (src/ell.c:46-73).  On a structured
// grid the resulting column order of a row depends only on which of the
// (up to 8) elements around the row's node exist, i.e. on the node's
// boundary class, so it is computed here once by replaying that insertion
// order on the local element patch.
#include <cstring>

#include "mm_internal.h"

namespace mm {

void build_slot_tables(SlotTables *t)
{
  std::memset(t, 0, sizeof(*t));
  for (int cz = 0; cz < 3; cz++)
    for (int cy = 0; cy < 3; cy++)
      for (int cx = 0; cx < 3; cx++) {
        const int cls = cx + 3 * cy + 9 * cz;
        const int cc[3] = {cx, cy, cz};
        // element offsets present along each axis, ascending
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
          lo[d] = (cc[d] == 0) ? 0 : -1;  // low face: no element below
          hi[d] = (cc[d] == 2) ? -1 : 0;  // high face: no element above
        }
        int nslot = 0, nelem = 0;
        int slot_of[27];
        for (int i = 0; i < 27; i++) slot_of[i] = -1;
        // ascending element id = z-major, then y, then x
        for (int oz = lo[2]; oz <= hi[2]; oz++)
          for (int oy = lo[1]; oy <= hi[1]; oy++)
            for (int ox = lo[0]; ox <= hi[0]; ox++) {
              const int a = local_node(-ox, -oy, -oz);
              const int eo = (ox + 1) | ((oy + 1) << 1) | ((oz + 1) << 2);
              t->elem[cls][nelem++] = (uint8_t)(eo | (a << 3));
              for (int b = 0; b < 8; b++) {
                int bx, by, bz;
                local_coords(b, &bx, &by, &bz);
                const int key = (ox + bx + 1) + 3 * (oy + by + 1) + 9 * (oz + bz + 1);
                if (slot_of[key] < 0) {
                  slot_of[key] = nslot;
                  t->off[cls][nslot] = (uint8_t)key;
                  nslot++;
                }
                const int s = slot_of[key];
                t->contrib[cls][s][t->ncontrib[cls][s]++] = (uint16_t)(eo | (a << 3) | (b << 6));
              }
            }
        for (int key = 0; key < 27; key++) t->key_slot[cls][key] = (int8_t)slot_of[key];
        t->nslot[cls] = nslot;
        t->nelem[cls] = (uint8_t)nelem;
        t->diag_slot[cls] = (uint8_t)slot_of[13];
      }
}

}  // namespace mm
