/* Caller-side stand-in for BQCD's site numbering.  TEST INFRASTRUCTURE ONLY.
 *
 * QWS's loaders take BQCD-ordered arrays and ask BQCD for each site's index within
 * its parity (std_xyzt2i_, 1-based) and its parity (e_o_) (qws_bqcd.cc:56-57).  BQCD
 * is not part of QWS; the tests load this library into the global symbol scope and
 * lay their arrays out with oracle/bqcd.py, which numbers the sites the same way:
 *   parity  = (x + y + z + t) % 2, the parity QWS's even/odd kernels assume (the
 *             parity-pe site at compressed x index xh has x = 2 xh + (pe+y+z+t)%2,
 *             deo_in_d.cc:80-85);
 *   index   = volh - lex/2 with lex = x + Lx (y + Ly (z + Lz t)): each parity in
 *             reverse lexicographic order, so a loader that ignored the index would
 *             put every site in the wrong place. */
static int L[4];

void bqcd_sites_set_(const int* l) {
  for (int i = 0; i < 4; ++i) L[i] = l[i];
}

int std_xyzt2i_(int* j) {
  const long lex = j[0] + (long)L[0] * (j[1] + (long)L[1] * (j[2] + (long)L[2] * j[3]));
  const long volh = (long)L[0] * L[1] * L[2] * L[3] / 2;
  return (int)(volh - lex / 2);
}

int e_o_(int* j) { return (j[0] + j[1] + j[2] + j[3]) & 1; }
