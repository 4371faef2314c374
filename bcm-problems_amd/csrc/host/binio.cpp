// Binary matrix and vector files in the reference's format: the raw
// native-endian int and double arrays CSMatrixRS::WriteMtrBin
// (smatrixrs.cpp:3210-3227) and CSVector::WriteVectBin (svector.cpp:537-546)
// write, one after another with no header or padding, so that files move
// between a Bslib program and this library unchanged.
//
//   matrix: m, n, nsupr, nsupc, nlist, nzja, nza, blamx   (int)
//           sprndr[nsupr+1], sprndc[nsupc+1], list[nlist], ia[nlist+1],
//           ja[nzja], bsa[nzja]                          (int)
//           a[nza]                                        (double)
//   vector: nparts, nrhs, nv (int), vect[nv*nrhs] (double),
//           blksv[nparts+1], inda[nparts] (int)
//
// The reader accepts what the writer produces: square block matrices whose
// row list is the identity (every block row present, as after assembly) and
// whose blocks are stored in ja order (bsa ascending in steps of the block
// sizes, as CSMatrixRS keeps them); anything else is refused with an error.
#include <algorithm>
#include <cstdint>
#include <cstdio>

#include "bcm_host.hpp"

namespace bcm {

namespace {

struct File {
  std::FILE* f = nullptr;
  File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {
    if (!f) throw Error(kErrArg, "cannot open " + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  template <class T>
  void put(const T* p, size_t n) {
    if (n && std::fwrite(p, sizeof(T), n, f) != n) throw Error(kErrInternal, "short write");
  }
  template <class T>
  void get(T* p, size_t n) {
    if (n && std::fread(p, sizeof(T), n, f) != n) throw Error(kErrArg, "binary file ends early");
  }
  void close() {
    if (f && std::fclose(f) != 0) {
      f = nullptr;
      throw Error(kErrInternal, "close failed");
    }
    f = nullptr;
  }
};

}  // namespace

void write_mtr_bin(const BlockCsr& a, const std::string& path) {
  if (a.nza() > INT32_MAX) throw Error(kErrArg, "WriteMtrBin: nza exceeds the format's int range");
  File f(path, "wb");
  const int n = a.n(), nsup = a.nsup, nzja = a.nzja(), nza = (int)a.nza();
  int blamx = 0;
  for (int i = 0; i < nsup; ++i) blamx = std::max(blamx, a.bla(i));
  const int head[8] = {n, n, nsup, nsup, nsup, nzja, nza, blamx};
  f.put(head, 8);
  f.put(a.sprnd.data(), nsup + 1);
  f.put(a.sprnd.data(), nsup + 1);
  std::vector<int> list(nsup);
  for (int i = 0; i < nsup; ++i) list[i] = i;
  f.put(list.data(), nsup);
  f.put(a.ia.data(), nsup + 1);
  f.put(a.ja.data(), nzja);
  std::vector<int> bsa(nzja);
  for (int j = 0; j < nzja; ++j) bsa[j] = (int)a.bsa[j];
  f.put(bsa.data(), nzja);
  f.put(a.a.data(), a.a.size());
  f.close();
}

BlockCsr read_mtr_bin(const std::string& path) {
  File f(path, "rb");
  int head[8];
  f.get(head, 8);
  const int m = head[0], n = head[1], nsupr = head[2], nsupc = head[3], nlist = head[4], nzja = head[5],
            nza = head[6];
  if (m != n || nsupr != nsupc || nlist != nsupr || nsupr < 0 || nzja < 0 || nza < 0)
    throw Error(kErrArg, "ReadMtrBin: only square block matrices with every block row listed are supported");
  BlockCsr a;
  a.nsup = nsupr;
  a.sprnd.resize(nsupr + 1);
  f.get(a.sprnd.data(), nsupr + 1);
  std::vector<int> sprndc(nsupc + 1), list(nlist);
  f.get(sprndc.data(), nsupc + 1);
  f.get(list.data(), nlist);
  if (sprndc != a.sprnd) throw Error(kErrArg, "ReadMtrBin: row and column partitions differ");
  for (int i = 0; i < nlist; ++i)
    if (list[i] != i) throw Error(kErrArg, "ReadMtrBin: the row list is not the identity");
  a.ia.resize(nsupr + 1);
  f.get(a.ia.data(), nsupr + 1);
  if (a.ia[0] != 0 || a.ia[nsupr] != nzja) throw Error(kErrArg, "ReadMtrBin: inconsistent row pointers");
  a.ja.resize(nzja);
  f.get(a.ja.data(), nzja);
  std::vector<int> bsa(nzja);
  f.get(bsa.data(), nzja);
  a.a.resize(nza);
  f.get(a.a.data(), nza);
  // blocks must lie in ja order, back to back
  a.bsa.resize(nzja);
  int64_t off = 0;
  for (int i = 0; i < nsupr; ++i)
    for (int j = a.ia[i]; j < a.ia[i + 1]; ++j) {
      if (a.ja[j] < 0 || a.ja[j] >= nsupr) throw Error(kErrArg, "ReadMtrBin: block column out of range");
      if (bsa[j] != off) throw Error(kErrArg, "ReadMtrBin: blocks are not stored in ja order");
      a.bsa[j] = off;
      off += (int64_t)a.bla(i) * a.bla(a.ja[j]);
    }
  if (off != nza) throw Error(kErrArg, "ReadMtrBin: nza does not match the block sizes");
  a.set_blamx();
  return a;
}

void write_vect_bin(const std::string& path, int n, int nrhs, const double* v) {
  File f(path, "wb");
  const int head[3] = {1, nrhs, n};
  f.put(head, 3);
  f.put(v, (size_t)n * nrhs);
  const int blksv[2] = {0, n}, inda[1] = {0};
  f.put(blksv, 2);
  f.put(inda, 1);
  f.close();
}

void read_vect_bin(const std::string& path, int* n, int* nrhs, std::vector<double>* v) {
  File f(path, "rb");
  int head[3];
  f.get(head, 3);
  const int nparts = head[0];
  if (nparts < 0 || head[1] < 0 || head[2] < 0) throw Error(kErrArg, "ReadVectBin: negative sizes");
  *nrhs = head[1];
  *n = head[2];
  if (v) {
    v->resize((size_t)head[1] * head[2]);
    f.get(v->data(), v->size());
    std::vector<int> tail((size_t)2 * nparts + 1);
    f.get(tail.data(), tail.size());  // blksv, inda: read to check the file is whole
  }
}

}  // namespace bcm
