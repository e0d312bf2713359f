/*
 * uni10.hpp - Uni10 C++ API (Qnum, Bond, Matrix, UniTensor, Network) over the
 * MI355X kernel library (tn_hip.h).
 *
 * Same class names, signatures and argument meanings as Uni10 v1.0
 * (uni10/src/uni10/datatype/Qnum.h, data-structure/Bond.h, Block.h,
 * tensor-network/Matrix.h, UniTensor.h, Network.h) for the real (FP64) path,
 * so caller code such as the iTEBD example compiles unchanged.  Differences:
 *   - element buffers always live in device memory (HBM); isOngpu() is true;
 *     host access (operator[], at, getRawElem, getHostElem) copies back;
 *   - errors throw std::runtime_error with the reference's message text;
 *   - a thread-local "current stream" (setStream) selects where work is queued.
 */
#ifndef UNI10_AMD_UNI10_HPP
#define UNI10_AMD_UNI10_HPP

#include <complex>
#include <cstddef>
#include <cstdint>
#include <iosfwd>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace uni10 {

typedef double Real;

enum bondType { BD_IN = 1, BD_OUT = -1 };
// datatype flags (datatype.hpp:90-99): RTYPE real FP64, CTYPE complex FP64
enum rflag { RNULL = 0, RTYPE = 1 };
enum cflag { CNULL = 0, CTYPE = 2 };
enum parityType { PRT_EVEN = 0, PRT_ODD = 1 };
enum parityFType { PRTF_EVEN = 0, PRTF_ODD = 1 };

/* GEMM arithmetic for contractions (storage stays FP64). */
enum GemmMode { GEMM_F64 = 0, GEMM_F32 = 1, GEMM_BF16X3 = 2, GEMM_BF16X6 = 3, GEMM_BF16_EXACT = 4 };
void setGemmMode(GemmMode m);          // thread-local; default GEMM_F64
GemmMode gemmMode();
void setAutotune(bool on);             // thread-local; time tile shapes on first use (default on)
void setStream(void* stream);          // thread-local; NULL = library default
void* currentStream();
void synchronize();                    // wait for the current stream

// ---------------------------------------------------------------------------
// Qnum (Qnum.h:52-170)
class Qnum {
 public:
  static const int U1_UPB = 1000;
  static const int U1_LOB = -1000;
  Qnum(int U1 = 0, parityType prt = PRT_EVEN);
  Qnum(parityFType prtF, int U1 = 0, parityType prt = PRT_EVEN);
  int U1() const { return m_U1; }
  parityType prt() const { return m_prt; }
  parityFType prtF() const { return m_prtF; }
  void assign(int U1 = 0, parityType prt = PRT_EVEN);
  void assign(parityFType prtF, int U1 = 0, parityType prt = PRT_EVEN);
  static bool isFermionic();
  long int hash() const { return (long int)m_U1 * 4 + m_prt * 2 + m_prtF; }
  friend bool operator<(const Qnum& a, const Qnum& b) { return a.hash() < b.hash(); }
  friend bool operator<=(const Qnum& a, const Qnum& b) { return a.hash() <= b.hash(); }
  friend bool operator==(const Qnum& a, const Qnum& b) {
    return a.m_U1 == b.m_U1 && a.m_prt == b.m_prt && a.m_prtF == b.m_prtF;
  }
  friend bool operator!=(const Qnum& a, const Qnum& b) { return !(a == b); }
  friend Qnum operator-(const Qnum& q);
  friend Qnum operator*(const Qnum& a, const Qnum& b);
  friend std::ostream& operator<<(std::ostream& os, const Qnum& q);

 private:
  int m_U1;
  parityType m_prt;
  parityFType m_prtF;
};

// ---------------------------------------------------------------------------
// Bond (Bond.h:51-180)
class Bond {
 public:
  Bond() : m_type(BD_IN), m_dim(0) {}
  Bond(bondType tp, size_t dim);
  Bond(bondType tp, const std::vector<Qnum>& qnums);
  void assign(bondType tp, size_t dim);
  void assign(bondType tp, const std::vector<Qnum>& qnums);
  bondType type() const { return m_type; }
  int dim() const { return m_dim; }
  std::map<Qnum, int> degeneracy() const;
  std::vector<Qnum> Qlist() const;
  Bond& change(bondType tp);
  Bond& dummy_change(bondType tp);
  Bond& combine(Bond bd);
  friend bool operator==(const Bond& a, const Bond& b);
  friend std::ostream& operator<<(std::ostream& os, const Bond& b);
  // run-length representation (public read access for planners)
  const std::vector<Qnum>& qnums() const { return Qnums; }
  const std::vector<int>& qdegs() const { return Qdegs; }
  const std::vector<int>& qoffsets() const { return offsets; }

 private:
  void setting(const std::vector<Qnum>& qnums);
  bondType m_type;
  int m_dim;
  std::vector<Qnum> Qnums;
  std::vector<int> Qdegs;
  std::vector<int> offsets;
  friend class UniTensor;
};
Bond combine(bondType tp, const std::vector<Bond>& bds);
Bond combine(const std::vector<Bond>& bds);

// ---------------------------------------------------------------------------
// Device buffer (owning, pool-backed)
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n_doubles, bool zero = false);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept;
  DevBuf& operator=(DevBuf&& o) noexcept;
  double* get() const { return p_; }
  size_t size() const { return n_; }
  void reset(size_t n_doubles, bool zero = false);

 private:
  double* p_ = nullptr;
  size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// Matrix (Matrix.h:47-391): owning dense or diagonal FP64 matrix in HBM.
class Matrix {
 public:
  Matrix();
  Matrix(size_t Rnum, size_t Cnum, bool diag = false, bool ongpu = true);
  Matrix(size_t Rnum, size_t Cnum, const double* elem, bool diag = false, bool src_ongpu = false);
  // complex matrices (MatrixComplex.cpp): elements interleaved (re, im) in HBM
  Matrix(cflag tp, size_t Rnum, size_t Cnum, bool diag = false);
  Matrix(cflag tp, size_t Rnum, size_t Cnum, const std::complex<double>* elem, bool diag = false);
  Matrix(const Matrix& m);
  Matrix(Matrix&& m) noexcept;
  Matrix& operator=(const Matrix& m);
  Matrix& operator=(Matrix&& m) noexcept;
  ~Matrix() = default;

  size_t row() const { return Rnum; }
  size_t col() const { return Cnum; }
  bool isDiag() const { return diag; }
  bool isOngpu() const { return true; }
  size_t elemNum() const { return diag ? (Rnum < Cnum ? Rnum : Cnum) : Rnum * Cnum; }
  int typeID() const { return ctype ? 2 : 1; }
  double* getElem() const { return buf.get(); }     // device pointer (interleaved for CTYPE)
  std::vector<double> getHostElem() const;          // CTYPE: interleaved (re, im)
  std::vector<std::complex<double>> getHostElemC() const;
  void setElem(const double* elem, bool src_ongpu = false);
  void setElem(const std::vector<double>& elem, bool src_ongpu = false);
  double operator[](size_t idx) const;
  double at(size_t r, size_t c) const;
  void set_zero();
  void identity();
  void randomize(uint64_t seed = 0);                 // SplitMix64 uniform[-1,1)
  Matrix& resize(size_t row, size_t col);
  Matrix& transpose();
  Matrix& cTranspose() { return transpose(); }
  double norm() const;
  double sum() const;
  double trace() const;
  Matrix& operator*=(double a);
  Matrix& operator+=(const Matrix& m);
  // CTYPE: U, S (real diagonal), VT with A = U S VT; eigh of a Hermitian
  // matrix: real eigenvalues, eigenvector rows E with A = E^H diag E
  std::vector<Matrix> svd() const;
  std::vector<Matrix> eigh() const;
  Matrix getDiag() const;
  friend Matrix operator*(const Matrix& a, const Matrix& b);
  friend Matrix operator*(const Matrix& a, double s);
  friend Matrix operator*(double s, const Matrix& a);
  friend Matrix operator+(const Matrix& a, const Matrix& b);
  friend bool operator==(const Matrix& a, const Matrix& b);
  friend std::ostream& operator<<(std::ostream& os, const Matrix& m);

 private:
  size_t Rnum = 0, Cnum = 0;
  bool diag = false;
  bool ctype = false;
  DevBuf buf;
  size_t nstore() const { return elemNum() * (ctype ? 2 : 1); }   // doubles held
  void real_only(const char* what) const;
  // storage left unset, for results a kernel writes in full
  struct uninit_t {};
  Matrix(uninit_t, size_t Rnum, size_t Cnum, bool diag);
  friend class UniTensor;
  friend class ITebd;
};
typedef Matrix Block;  // read-only views are returned as copies
Matrix takeExp(double a, const Matrix& mat);
Matrix exph(double a, const Matrix& mat);
Matrix RDotR(const Matrix& a, const Matrix& b);
// lanczosEigh (Block.h:389, Block.cpp:474-506): lowest eigenvalue E0 of a
// square matrix by Lanczos; psi is the start vector on entry and the ground
// state on return.  Returns the steps taken; throws when max_iter steps do
// not converge.
size_t lanczosEigh(Matrix& ori_mat, double& E0, Matrix& psi, size_t max_iter = 1000, double err_tol = 5e-15);

// A swap gate between bonds b1 and b2 of one tensor (uni10_tools.h: _Swap).
struct _Swap {
  int b1;
  int b2;
};

// ---------------------------------------------------------------------------
// UniTensor (UniTensor.h:58-971), real FP64.
struct BlockInfo {
  size_t Rnum = 0, Cnum = 0, offset = 0;
};
struct Layout;                                   // block plan (grouping result)
typedef std::shared_ptr<const Layout> LayoutPtr;

class UniTensor {
 public:
  UniTensor();
  explicit UniTensor(double val);
  UniTensor(const std::vector<Bond>& bonds, const std::string& name = "");
  UniTensor(const std::vector<Bond>& bonds, std::vector<int>& labels, const std::string& name = "");
  // complex tensors (UniTensor.h:798-800, UniTensorComplex.cpp)
  UniTensor(cflag tp, const std::vector<Bond>& bonds, const std::string& name = "");
  UniTensor(cflag tp, const std::vector<Bond>& bonds, std::vector<int>& labels, const std::string& name = "");
  explicit UniTensor(const std::string& fname);   // binary dump (current or older format)
  UniTensor(const UniTensor& T);
  UniTensor(UniTensor&& T) noexcept;
  UniTensor& operator=(const UniTensor& T);
  UniTensor& operator=(UniTensor&& T) noexcept;
  ~UniTensor();

  UniTensor& assign(const std::vector<Bond>& bonds);
  void setName(const std::string& name) { m_name = name; }
  std::string getName() const { return m_name; }
  void setLabel(const std::vector<int>& newLabels);
  void setLabel(int* newLabels);
  // (a complex tensor keeps its (re, im) pair as a hidden last bond; these
  // accessors show the tensor's own bonds)
  std::vector<int> label() const {
    return m_ctype ? std::vector<int>(labels.begin(), labels.end() - 1) : labels;
  }
  int label(size_t idx) const { return label().at(idx); }
  std::vector<Bond> bond() const {
    return m_ctype ? std::vector<Bond>(bonds.begin(), bonds.end() - 1) : bonds;
  }
  Bond bond(size_t idx) const { return bond().at(idx); }
  size_t bondNum() const { return bonds.size() - (m_ctype ? 1 : 0); }
  size_t inBondNum() const { return RBondNum; }
  size_t elemNum() const { return m_ctype ? m_elemNum / 2 : m_elemNum; }
  size_t blockNum() const;
  std::vector<Qnum> blockQnum() const;
  Qnum blockQnum(size_t idx) const;
  int typeID() const { return m_ctype ? 2 : 1; }

  void setRawElem(const std::vector<double>& raw);
  void setRawElem(const double* raw);
  Matrix getRawElem() const;                         // CTYPE: a complex matrix
  void setRawElem(const std::vector<std::complex<double>>& raw);
  void setRawElem(const std::complex<double>* raw);
  void setElem(const std::complex<double>* elem, bool src_ongpu = false);
  std::complex<double> at(cflag tp, size_t idx) const;
  void setElem(const double* elem, bool src_ongpu = false);
  void setElem(const std::vector<double>& elem, bool src_ongpu = false);
  double* getElem() const { return buf.get(); }     // device pointer, block layout
  std::vector<double> getHostElem() const;
  double at(const std::vector<size_t>& idxs) const;
  double at(size_t idx) const;
  double operator[](size_t idx) const { return at(idx); }

  void putBlock(const Matrix& mat);
  void putBlock(const Qnum& qnum, const Matrix& mat);
  Matrix getBlock(bool diag = false) const;
  Matrix getBlock(const Qnum& qnum, bool diag = false) const;
  std::map<Qnum, Matrix> getBlocks() const;

  UniTensor& permute(const std::vector<int>& newLabels, int inBondNum);
  UniTensor& permute(int* newLabels, int inBondNum);
  UniTensor& permute(int inBondNum);
  UniTensor& transpose();
  UniTensor& operator*=(double a);
  double norm() const;
  UniTensor& normalize();
  void randomize(uint64_t seed = 0);
  void set_zero();
  void save(const std::string& fname) const;
  std::string printDiagram() const;
  // Fermionic swap gates (UniTensor.h:664-665): the bond pairs (shared with
  // Tb, not shared) that cross when this tensor passes Tb, and the in-place
  // sign (-1)^(sum over gates of prtF(b1) * prtF(b2)) per Qidx sub-block.
  std::vector<_Swap> exSwap(const UniTensor& Tb) const;
  void addGate(const std::vector<_Swap>& swaps);

  friend UniTensor contract(UniTensor& Ta, UniTensor& Tb, bool fast);
  friend UniTensor contract_impl(UniTensor& Ta, UniTensor& Tb, bool fast, bool allow_gather);
  friend UniTensor operator*(const UniTensor& Ta, const UniTensor& Tb);
  friend UniTensor operator*(const UniTensor& T, double a);
  friend UniTensor operator*(double a, const UniTensor& T);
  friend UniTensor otimes(const UniTensor& Ta, const UniTensor& Tb);
  friend std::ostream& operator<<(std::ostream& os, const UniTensor& T);
  friend class Network;

  // Block plan access (for planners / bindings)
  const std::map<Qnum, BlockInfo>& blockInfo() const;
  LayoutPtr layout() const { return m_layout; }
  bool hasElem() const { return (status & HAVEELEM) != 0; }
  bool hasBond() const { return (status & HAVEBOND) != 0; }

  // profiling counters (UniTensor.h:945-948)
  static std::string profile(bool print = true);

  // Construction without the zero fill, for outputs that are fully written.
  enum InitFill { Zero, NoZero };
  UniTensor(const std::vector<Bond>& bonds, const std::string& name, InitFill fill);

 private:
  void initUniT(bool zero = true);
  void make_complex(const std::vector<Bond>& b, bool zero);
  void real_only(const char* what) const;
  UniTensor& permute_full(const std::vector<int>& newLabels, int inBondNum);
  friend UniTensor contract_complex(UniTensor& Ta, UniTensor& Tb, bool fast, bool allow_gather);
  bool m_ctype = false;
  std::string m_name;
  std::vector<Bond> bonds;
  std::vector<int> labels;
  LayoutPtr m_layout;
  DevBuf buf;
  size_t m_elemNum = 0;
  int RBondNum = 0;
  int status = 0;
  static const int HAVEBOND = 1;
  static const int HAVEELEM = 2;
  static int64_t COUNTER, ELEMNUM, MAXELEMNUM, MAXELEMTEN;
};
UniTensor contract(UniTensor& Ta, UniTensor& Tb, bool fast = false);
UniTensor otimes(const UniTensor& Ta, const UniTensor& Tb);

// ---------------------------------------------------------------------------
// Network (Network.h:52-230)
struct NetNode;
class Network {
 public:
  explicit Network(const std::string& fname);
  Network(const std::string& fname, const std::vector<UniTensor*>& tens);
  ~Network();
  Network(const Network&) = delete;
  Network& operator=(const Network&) = delete;
  void putTensor(size_t idx, const UniTensor* T, bool force = true);
  void putTensor(size_t idx, const UniTensor& T, bool force = true);
  void putTensor(const std::string& name, const UniTensor* T, bool force = true);
  void putTensor(const std::string& name, const UniTensor& T, bool force = true);
  void putTensorT(const std::string& nameT, const UniTensor* T, bool force = true);
  void putTensorT(const std::string& nameT, const UniTensor& T, bool force = true);
  UniTensor launch(const std::string& name = "");
  // Graph replay (no counterpart in the reference): with on, launch()
  // captures the whole contraction (copies of the leaf tensors, every
  // permute and sector GEMM, the final permute) into a HIP graph once per
  // leaf layout and buffer, then replays it; the leaves keep their put
  // state and the result is a copy of the graph's output.
  void setGraph(bool on);
  std::string profile(bool print = true);
  std::string tree() const;     // contraction tree as nested (a b) of tensor names
  friend std::ostream& operator<<(std::ostream& os, Network& net);

 private:
  void fromfile(const std::string& fname);
  void construct();
  void destruct();
  int rollcall();
  void preprint(std::ostream& os, const NetNode* nd, int layer) const;
  void findConOrd(NetNode* nd);
  void addSwap();
  UniTensor merge_on(NetNode* nd, std::vector<UniTensor>& W);
  UniTensor run_on_copies();
  UniTensor launch_graph(const std::string& name);
  void drop_graph();
  void matching(NetNode* sbj, NetNode* tar);
  void branch(NetNode* sbj, NetNode* tar);
  UniTensor merge(NetNode* nd);
  std::vector<std::string> names;
  std::map<std::string, size_t> name2pos;
  std::vector<std::vector<int>> label_arr;
  std::vector<int> Rnums;
  std::vector<int> order;
  std::vector<int> brakets;
  std::vector<UniTensor*> tensors;
  std::vector<NetNode*> leafs;
  std::vector<NetNode*> owned;   // internal nodes
  std::vector<int> conOrder;                  // leaves in pre-order of the tree
  std::vector<bool> swapflags;                // gates applied to tensors[t]
  std::vector<std::vector<_Swap>> swaps_arr;  // gates of tensors[t]
  bool use_graph = false;
  void* graph = nullptr;                      // captured launch
  std::string graph_sig, graph_failed_sig;    // leaf layouts/buffers it was captured for
  UniTensor* graph_out = nullptr;             // its output (graph-owned memory)
  NetNode* root = nullptr;
  bool load = false;
};

}  // namespace uni10
#endif  // UNI10_AMD_UNI10_HPP
