// Network: `.net` parser, contraction-tree planner and post-order executor.
// Follows uni10/src/uni10/tensor-network/lib/Network.cpp: Node::contract /
// metric / cal_elemNum (:63-219), fromfile (:271-426), construct (:428-506),
// putTensor (:508-574), putTensorT (:576-591), branch/matching (:602-670),
// launch/merge (:697-739).  Every pairwise contraction runs through
// contract() and so through the grouped GEMM and permute kernels.
#include <algorithm>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>

#include "common.h"
#include "layout.h"

namespace uni10 {

struct NetNode {
  UniTensor* T = nullptr;     // leaf if non-null
  std::vector<int> labels;
  std::vector<Bond> bonds;
  int64_t elemNum = 0;
  std::string name;
  NetNode* parent = nullptr;
  NetNode* left = nullptr;
  NetNode* right = nullptr;
  float point = 0;
};

static int64_t cal_elem_num(const std::vector<Bond>& bonds) {
  try {
    return (int64_t)get_layout(bonds)->elemNum;
  } catch (const std::exception&) {
    return 0;   // no common charge: the reference's sum over no blocks
  }
}

// Bonds and labels of the contraction of a with b (Node::contract); returns
// the number of contracted bonds.
static int merge_bonds(const NetNode* a, const NetNode* b, std::vector<Bond>& cb, std::vector<int>& cl) {
  const int na = (int)a->bonds.size(), nbb = (int)b->bonds.size();
  std::vector<int> markB(nbb, 0);
  int con = 0;
  cb.clear();
  cl.clear();
  for (int i = 0; i < na; ++i) {
    bool match = false;
    for (int j = 0; j < nbb; ++j)
      if (a->labels[i] == b->labels[j]) {
        markB[j] = 1;
        match = true;
        con++;
        break;
      }
    if (!match) {
      cl.push_back(a->labels[i]);
      cb.push_back(a->bonds[i]);
    }
  }
  for (int j = 0; j < nbb; ++j)
    if (!markB[j]) {
      cl.push_back(b->labels[j]);
      cb.push_back(b->bonds[j]);
    }
  const int r = na - con;
  for (int i = 0; i < (int)cb.size(); ++i) cb[i].change(i < r ? BD_IN : BD_OUT);
  return con;
}

static float node_metric(const NetNode* a, const NetNode* b) {
  std::vector<Bond> cb;
  std::vector<int> cl;
  int con = merge_bonds(a, b, cb, cl);
  if (con == 0) return -1.0f;
  int64_t ne = cal_elem_num(cb);
  return float(a->elemNum + b->elemNum) / ne;
}

// ---------------------------------------------------------------------------
static void trim(std::string& s) {
  size_t a = s.find_first_not_of(" \t\r");
  size_t b = s.find_last_not_of(" \t\r");
  s = (a == std::string::npos) ? std::string() : s.substr(a, b - a + 1);
}

void Network::fromfile(const std::string& fname) {
  std::ifstream in(fname.c_str());
  if (!in.is_open()) detail::fail("Error in opening file '" + fname + "'.");
  const std::string tar("1234567890-");
  std::vector<std::string> ord;
  std::string str;
  int lnum = 0;
  while (lnum < 1000) {
    std::getline(in, str);
    if (in.eof()) break;   // a last line without a newline is not read (reference behaviour)
    size_t pos = str.find(':');
    if (pos == std::string::npos) break;
    std::string name = str.substr(0, pos);
    trim(name);
    if (name == "ORDER") {
      if (str.find('(', pos + 1) == std::string::npos) {
        const std::string del(" ,;");
        size_t p = pos;
        while ((p = str.find_first_not_of(del, p + 1)) != std::string::npos) {
          size_t e = str.find_first_of(del, p + 1);
          ord.push_back(e == std::string::npos ? str.substr(p) : str.substr(p, e - p));
          p = e;
          if (p == std::string::npos) break;
        }
      } else {
        const std::string del(" ,;*()");
        std::string pending;
        size_t i = pos + 1;
        while (i < str.size()) {
          if (del.find(str[i]) != std::string::npos) {
            pending += str[i++];
            continue;
          }
          size_t j = i;
          while (j < str.size() && del.find(str[j]) == std::string::npos) ++j;
          int closes = (int)std::count(pending.begin(), pending.end(), ')');
          int opens = (int)std::count(pending.begin(), pending.end(), '(');
          if (closes) brakets.push_back(-closes);
          if (opens) brakets.push_back(opens);
          ord.push_back(str.substr(i, j - i));
          brakets.push_back(0);
          pending.clear();
          i = j;
        }
        int closes = (int)std::count(pending.begin(), pending.end(), ')');
        if (closes) brakets.push_back(-closes);
        int sum = 0;
        for (int b : brakets) sum += b;
        if (sum != 0)
          detail::fail("Error in the network file '" + fname +
                       "'. There are imbalance brackets when specifying the contraction order.");
      }
      break;
    }
    name2pos[name] = names.size();
    names.push_back(name);
    std::vector<int> labels;
    int Rnum = -1;
    bool semi = false;
    size_t i = pos + 1;
    while (i < str.size()) {
      char ch = str[i];
      if (ch == ';') semi = true;
      if (tar.find(ch) != std::string::npos) {
        size_t j = i;
        while (j < str.size() && tar.find(str[j]) != std::string::npos) ++j;
        if (Rnum == -1 && semi) Rnum = (int)labels.size();
        labels.push_back((int)std::strtol(str.substr(i, j - i).c_str(), nullptr, 10));
        i = j;
        continue;
      }
      ++i;
    }
    label_arr.push_back(labels);
    Rnums.push_back(Rnum == -1 ? (int)labels.size() : Rnum);
    lnum++;
  }
  const int numT = (int)names.size() - 1;
  if (numT < 0 || names[numT] != "TOUT")
    detail::fail("Error in the network file '" + fname +
                 "'. Missing TOUT tensor. One must specify the bond labels for the resulting tensor by giving TOUT tag.");
  if (names.size() <= 2)
    detail::fail("Error in the network file '" + fname + "'. There must be at least two tensors in a tensor network.");
  order.assign(numT, 0);
  if (!ord.empty()) {
    if ((int)ord.size() != numT)
      detail::fail("Error in the network file '" + fname + "'. Some tensors are missing in the contraction order.");
    std::vector<bool> found(numT, false);
    for (int i = 0; i < numT; ++i) {
      auto it = name2pos.find(ord[i]);
      if (it == name2pos.end())
        detail::fail("Error in the network file '" + fname + "'. '" + ord[i] +
                     "' in the contraction order is not in the list of tensors above.");
      order[i] = (int)it->second;
      if (found[order[i]])
        detail::fail("Error in the network file '" + fname + "'. '" + ord[i] +
                     "' appears more than once in the contraction order.");
      found[order[i]] = true;
    }
  } else {
    for (int i = 0; i < numT; ++i) order[i] = i;
  }
}

Network::Network(const std::string& fname) {
  fromfile(fname);
  const int n = (int)label_arr.size() - 1;
  leafs.assign(n, nullptr);
  tensors.assign(n, nullptr);
  swapflags.assign(n, false);
  swaps_arr.assign(n, std::vector<_Swap>());
}

Network::Network(const std::string& fname, const std::vector<UniTensor*>& tens) {
  fromfile(fname);
  const int n = (int)label_arr.size() - 1;
  if ((int)tens.size() != n)
    detail::fail("The size of the input vector of tensors does not match for the number of tensors in the network file '" +
                 fname + "'.");
  leafs.assign(n, nullptr);
  tensors.assign(n, nullptr);
  swapflags.assign(n, false);
  swaps_arr.assign(n, std::vector<_Swap>());
  for (int i = 0; i < n; ++i) putTensor((size_t)i, tens[i]);
  construct();
}

Network::~Network() {
  drop_graph();
  for (NetNode* p : owned) delete p;
  for (NetNode* p : leafs) delete p;
  for (UniTensor* t : tensors) delete t;
}

void Network::putTensor(size_t idx, const UniTensor* T, bool force) {
  if (T && T->typeID() != 1) detail::fail("Network: complex tensors are not provided (contract them with contract())");
  if (idx >= label_arr.size() - 1)
    detail::fail("Index exceeds the number of the tensors in the list of network file.");
  if (!force && load) destruct();
  if ((int)T->inBondNum() != Rnums[idx])
    detail::fail("The number of in-coming bonds does not match with the tensor '" + names[idx] +
                 "' specified in network file");
  if (leafs[idx] != nullptr) {
    *tensors[idx] = *T;
    tensors[idx]->setLabel(label_arr[idx]);
    tensors[idx]->setName(names[idx]);
    swapflags[idx] = false;
  } else {
    if (!T->hasBond()) detail::fail("Cannot create node of a network from tensor without bond.");
    UniTensor* t = new UniTensor(*T);
    t->setName(names[idx]);
    t->setLabel(label_arr[idx]);
    tensors[idx] = t;
    NetNode* nd = new NetNode();
    nd->T = t;
    nd->labels = t->label();
    nd->bonds = t->bond();
    nd->elemNum = (int64_t)t->elemNum();
    nd->name = names[idx];
    leafs[idx] = nd;
  }
}

void Network::putTensor(size_t idx, const UniTensor& T, bool force) { putTensor(idx, &T, force); }

void Network::putTensor(const std::string& name, const UniTensor* T, bool force) {
  auto it = name2pos.find(name);
  if (it == name2pos.end()) detail::fail("There is no tensor named '" + name + "' in the network file");
  putTensor(it->second, T, force);
}

void Network::putTensor(const std::string& name, const UniTensor& T, bool force) { putTensor(name, &T, force); }

void Network::putTensorT(const std::string& nameT, const UniTensor* T, bool force) {
  auto it = name2pos.find(nameT);
  if (it == name2pos.end()) detail::fail("There is no tensor named '" + nameT + "' in the network file");
  UniTensor t = *T;
  t.transpose();
  putTensor(it->second, &t, force);
}

void Network::putTensorT(const std::string& nameT, const UniTensor& T, bool force) { putTensorT(nameT, &T, force); }

// Network.cpp:679-695: the swap gates already applied are applied again,
// which undoes them (each sign squares to one).
void Network::destruct() {
  drop_graph();
  for (NetNode* p : owned) delete p;
  owned.clear();
  root = nullptr;
  for (NetNode* l : leafs)
    if (l) {
      l->parent = l->left = l->right = nullptr;
      l->point = 0;
    }
  conOrder.clear();
  for (size_t t = 0; t < tensors.size(); ++t) {
    if (Qnum::isFermionic() && swapflags[t]) {
      tensors[t]->addGate(swaps_arr[t]);
      swapflags[t] = false;
    }
    swaps_arr[t].clear();
  }
  load = false;
}

// findConOrd (Network.cpp:923-942): leaf tensor indices in pre-order.
void Network::findConOrd(NetNode* nd) {
  if (nd == nullptr || conOrder.size() == tensors.size()) return;
  if (nd->T) {
    auto it = std::find(tensors.begin(), tensors.end(), nd->T);
    if (it == tensors.end()) detail::fail("Fatal error(code = N2). Please contact the developer of the uni10 library.");
    conOrder.push_back((int)(it - tensors.begin()));
  }
  findConOrd(nd->left);
  findConOrd(nd->right);
}

// addSwap (Network.cpp:944-993): sorting the tensors from their order in the
// tree back to the file order, each time tensor b1 passes tensor b2 the
// bonds b1 shares with b2 cross the bonds it does not share; a bond pair
// crossed an even number of times cancels.
void Network::addSwap() {
  const int Tnum = (int)leafs.size();
  findConOrd(root);
  if ((int)conOrder.size() != Tnum)
    detail::fail("Fatal error(code = N3). Please contact the developer of the uni10 library.");
  std::vector<int> ord = conOrder;
  for (int i = 0; i < Tnum - 1; ++i)
    for (int j = 0; j < Tnum - i - 1; ++j)
      if (ord[j] > ord[j + 1]) {
        const int b1 = ord[j + 1], b2 = ord[j];
        std::vector<_Swap> sw = tensors[b1]->exSwap(*tensors[b2]);
        swaps_arr[b1].insert(swaps_arr[b1].end(), sw.begin(), sw.end());
        std::swap(ord[j], ord[j + 1]);
      }
  for (int t = 0; t < Tnum; ++t) {
    const int nb = (int)tensors[t]->bondNum();
    std::map<int, bool> recs;
    for (const _Swap& sp : swaps_arr[t]) {
      const int is = std::min(sp.b1, sp.b2), ib = std::max(sp.b1, sp.b2);
      auto it = recs.find(is * nb + ib);
      if (it != recs.end())
        it->second = !it->second;
      else
        recs[is * nb + ib] = true;
    }
    swaps_arr[t].clear();
    for (auto& kv : recs)
      if (kv.second) swaps_arr[t].push_back(_Swap{kv.first / nb, kv.first % nb});
  }
}

void Network::branch(NetNode* sbj, NetNode* tar) {
  NetNode* par = new NetNode();
  owned.push_back(par);
  merge_bonds(tar, sbj, par->bonds, par->labels);
  par->elemNum = cal_elem_num(par->bonds);
  if (sbj->parent == nullptr) {
    if (tar->parent != nullptr) {
      if (tar->parent->left == tar)
        tar->parent->left = par;
      else
        tar->parent->right = par;
      par->parent = tar->parent;
    } else {
      par->parent = nullptr;
      root = par;
    }
  } else {
    NetNode* old = tar->parent;
    if (old->parent != nullptr) {
      if (old->parent->left == old)
        old->parent->left = par;
      else
        old->parent->right = par;
      par->parent = old->parent;
    } else {
      par->parent = nullptr;
      root = par;
    }
    owned.erase(std::find(owned.begin(), owned.end(), old));
    delete old;
  }
  par->left = tar;
  par->right = sbj;
  tar->parent = par;
  sbj->parent = par;
  par->point = node_metric(tar, sbj);
  if (sbj->parent->parent != nullptr) {
    NetNode* s = sbj->parent;
    branch(s->parent->right, s->parent->left);
  }
}

void Network::matching(NetNode* sbj, NetNode* tar) {
  if (tar == nullptr) {
    root = sbj;
  } else if (tar->T == nullptr) {
    if (node_metric(sbj, tar) > 0) {
      if (!(tar->left != nullptr && tar->right != nullptr))
        detail::fail("Fatal error(code = N1). Please contact the developer of the uni10 library.");
      float tar_p = tar->point, lft_p = 0, rht_p = 0;
      if ((lft_p = node_metric(sbj, tar->left)) > tar_p || (rht_p = node_metric(sbj, tar->right)) > tar_p) {
        if (lft_p > rht_p)
          matching(sbj, tar->left);
        else
          matching(sbj, tar->right);
      } else {
        branch(sbj, tar);
      }
    } else {
      branch(sbj, tar);
    }
  } else {
    branch(sbj, tar);
  }
}

void Network::construct() {
  if (!brakets.empty()) {
    std::vector<NetNode*> stack;
    int cnt = 0;
    auto fold = [&]() {
      NetNode* par = new NetNode();
      owned.push_back(par);
      NetNode* l = stack[stack.size() - 2];
      NetNode* r = stack[stack.size() - 1];
      merge_bonds(l, r, par->bonds, par->labels);
      par->elemNum = cal_elem_num(par->bonds);
      par->left = l;
      par->right = r;
      l->parent = r->parent = par;
      stack.pop_back();
      stack.back() = par;
    };
    for (int b : brakets) {
      if (b < 0) {
        for (int c = 0; c < -b; ++c) {
          fold();
          if (stack.size() < 2) break;
        }
      } else if (b == 0) {
        NetNode* lf = leafs[order[cnt]];
        if (lf == nullptr)
          detail::fail("(((Tensor '" + names[order[cnt]] + "' has not yet been given.\n  Hint: Use addTensor() to add a tensor to a network.\n");
        stack.push_back(lf);
        cnt++;
      }
    }
    while (stack.size() > 1) fold();
    root = stack[0];
  } else {
    for (size_t i = 0; i < order.size(); ++i) {
      if (leafs[order[i]] == nullptr)
        detail::fail("Tensor '" + names[order[i]] + "' has not yet been given.\n  Hint: Use putTensor() to add a tensor to a network.\n");
      matching(leafs[order[i]], root);
    }
  }
  const std::vector<int>& tout = label_arr.back();
  if (root->labels.size() != tout.size())
    detail::fail("Error when constructing the network. The bond number of the resulting tensor is different from the bond number of 'TOUT'");
  for (int l : root->labels)
    if (std::find(tout.begin(), tout.end(), l) == tout.end())
      detail::fail("Error when constructing the network. The labels of the resulting tensor do not match with the labels of 'TOUT' in the network file");
  addSwap();
  load = true;
}

UniTensor Network::merge(NetNode* nd) {
  if (nd->left->T == nullptr) {
    UniTensor lt = merge(nd->left);
    if (nd->right->T == nullptr) {
      UniTensor rt = merge(nd->right);
      return contract_nopermute(lt, rt);
    }
    return contract_nopermute(lt, *nd->right->T);
  }
  if (nd->right->T == nullptr) {
    UniTensor rt = merge(nd->right);
    return contract_nopermute(*nd->left->T, rt);
  }
  return contract_nopermute(*nd->left->T, *nd->right->T);
}

void Network::setGraph(bool on) {
  use_graph = on;
  if (!on) drop_graph();
}

void Network::drop_graph() {
  delete graph_out;
  graph_out = nullptr;
  if (graph) tn_graph_destroy(graph);
  graph = nullptr;
  graph_sig.clear();
}

// merge() over working copies W of the leaves (W[t] for tensors[t]).
UniTensor Network::merge_on(NetNode* nd, std::vector<UniTensor>& W) {
  auto operand = [&](NetNode* n, UniTensor& tmp) -> UniTensor& {
    if (n->T == nullptr) {
      tmp = merge_on(n, W);
      return tmp;
    }
    const size_t t = std::find(leafs.begin(), leafs.end(), n) - leafs.begin();
    return W.at(t);
  };
  UniTensor lt, rt;
  UniTensor& a = operand(nd->left, lt);
  UniTensor& b = operand(nd->right, rt);
  return contract_nopermute(a, b);
}

UniTensor Network::run_on_copies() {
  std::vector<UniTensor> W;
  W.reserve(tensors.size());
  for (size_t t = 0; t < tensors.size(); ++t) {
    W.emplace_back(*tensors[t]);
    if (Qnum::isFermionic() && !swapflags[t]) W.back().addGate(swaps_arr[t]);
  }
  UniTensor T = merge_on(root, W);
  const std::vector<int>& tout = label_arr.back();
  if (tout.size() > 1) T.permute(tout, Rnums.back());
  return T;
}

UniTensor Network::launch_graph(const std::string& name) {
  std::ostringstream sig;
  for (size_t t = 0; t < tensors.size(); ++t) {
    const UniTensor& T = *tensors[t];
    sig << (T.layout() ? T.layout()->key : std::string()) << "@" << (const void*)T.getElem() << ":" << T.inBondNum();
    for (int l : T.label()) sig << "," << l;
    if (Qnum::isFermionic()) {
      sig << (swapflags[t] ? "g" : "G");
      for (const _Swap& w : swaps_arr[t]) sig << w.b1 << "-" << w.b2 << ";";
    }
    sig << "|";
  }
  const std::string key = sig.str();
  if (key == graph_failed_sig) {   // this leaf state could not be captured: run it directly
    UniTensor T = run_on_copies();
    T.setName(name);
    return T;
  }
  if (!graph || key != graph_sig) {
    drop_graph();
    {
      // every plan is built (and tuned) here: host uploads cannot be captured
      UniTensor warm = run_on_copies();
    }
    detail::check(tn_stream_sync(detail::stream()), "Network graph");
    detail::check(tn_graph_begin(detail::stream()), "Network graph");
    UniTensor* out = nullptr;
    try {
      out = new UniTensor(run_on_copies());
    } catch (const std::exception&) {
      void* g = nullptr;
      tn_graph_end(detail::stream(), &g);   // ends (and discards) the capture
      if (g) tn_graph_destroy(g);
      graph_failed_sig = key;
      UniTensor T = run_on_copies();
      T.setName(name);
      return T;
    }
    const int rc = tn_graph_end(detail::stream(), &graph);
    if (rc != TN_OK) {
      delete out;
      graph = nullptr;
      detail::check(rc, "Network graph");
    }
    graph_out = out;
    graph_sig = key;
  }
  detail::check(tn_graph_launch(graph, detail::stream()), "Network graph replay");
  UniTensor T(*graph_out);   // the next replay rewrites the graph's output
  T.setName(name);
  return T;
}

UniTensor Network::launch(const std::string& name) {
  if (!load) construct();
  if (use_graph) return launch_graph(name);
  for (size_t t = 0; t < tensors.size(); ++t)
    if (Qnum::isFermionic() && !swapflags[t]) {
      tensors[t]->addGate(swaps_arr[t]);
      swapflags[t] = true;
    }
  UniTensor T = merge(root);
  const std::vector<int>& tout = label_arr.back();
  if (tout.size() > 1) T.permute(tout, Rnums.back());
  T.setName(name);
  return T;
}

std::string Network::tree() const {
  std::function<std::string(const NetNode*)> rec = [&](const NetNode* n) -> std::string {
    if (n->T) return n->name;
    return "(" + rec(n->left) + " " + rec(n->right) + ")";
  };
  return root ? rec(root) : std::string();
}

std::string Network::profile(bool print) {
  if (!load) construct();
  int64_t total = 0, mx = 0;
  std::function<void(const NetNode*)> rec = [&](const NetNode* n) {
    if (!n) return;
    total += n->elemNum;
    mx = std::max(mx, n->elemNum);
    rec(n->left);
    rec(n->right);
  };
  rec(root);
  std::ostringstream os;
  os << "\n===== Network profile =====\n"
     << "Contraction tree: " << tree() << "\n"
     << "Memory requirement: " << total * (int64_t)sizeof(double) << " bytes\n"
     << "Largest tensor: " << mx << " elements\n"
     << "===========================\n";
  if (print) std::cout << os.str();
  return os.str();
}

int Network::rollcall() {
  if (!load) {
    for (size_t i = 0; i < leafs.size(); ++i)
      if (leafs[i] == nullptr) return (int)i;
    construct();
  }
  return -1;
}

// Network::preprint (Network.cpp:856-870): one line per node in pre-order,
// "name(elemNum): labels," indented by "|   " per level, "*" for inner nodes.
void Network::preprint(std::ostream& os, const NetNode* nd, int layer) const {
  if (nd == nullptr) return;
  for (int i = 0; i < layer; ++i) os << "|   ";
  if (nd->T)
    os << nd->name << "(" << nd->elemNum << "): ";
  else
    os << "*(" << nd->elemNum << "): ";
  for (int l : nd->labels) os << l << ", ";
  os << std::endl;
  preprint(os, nd->left, layer + 1);
  preprint(os, nd->right, layer + 1);
}

// operator<< (Network.cpp:872-907): the tensor list with in/out labels, then
// the contraction tree once every tensor has been given.
std::ostream& operator<<(std::ostream& os, Network& net) {
  os << std::endl;
  for (size_t i = 0; i < net.names.size(); ++i) {
    const int R = net.Rnums[i];
    const size_t nl = net.label_arr[i].size();
    os << net.names[i] << ": ";
    if (R) os << "i[";
    for (int l = 0; l < R; ++l) os << net.label_arr[i][l] << (l < R - 1 ? ", " : "");
    if (R) os << "] ";
    if (nl - R) os << "o[";
    for (size_t l = R; l < nl; ++l) os << net.label_arr[i][l] << (l < nl - 1 ? ", " : "");
    if (nl - R) os << "]";
    os << std::endl;
  }
  os << std::endl;
  if (net.rollcall() < 0)
    net.preprint(os, net.root, 0);
  else
    os << "\nSome tensors have not yet been given!\n\n";
  return os;
}

}  // namespace uni10
