/* Links against the static archive libconv_amd.a and calls the parts of the
 * C ABI that need no device: the version string and the size rules
 * (lc_extent).  Prints one line per check; exit status 0 when all hold. */
#include <stdio.h>
#include <string.h>

#include "libconv.h"

int main(void) {
  int bad = 0;
  const char *v = lc_version();
  printf("version %s\n", v);
  bad |= strstr(v, "gfx950") == NULL;
  /* magic filter grow +15, kinetic shrink input +28, sym8 grow +7 pairs */
  const int e[3] = {lc_extent(0, 1, 100, 1), lc_extent(2, -1, 100, 0), lc_extent(10, 1, 50, 1)};
  printf("extents %d %d %d\n", e[0], e[1], e[2]);
  bad |= e[0] != 115 || e[1] != 128 || e[2] != 57;
  printf("%s\n", bad ? "FAIL" : "OK");
  return bad;
}
