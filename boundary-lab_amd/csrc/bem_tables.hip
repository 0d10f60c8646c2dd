// Host-side mesh tables (no device code): vertex-sharing element pairs with their Duffy orientation codes, and
// the chunked slot tables of the regular assembly. Built per job before upload; these were the slowest part of
// a request's host work.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <string>
#include <vector>

#include "../../include/bemhip.h"
#include "bem_kernels.hpp"

int bem_fail_message(int code, const std::string& msg);  // bem_capi.hip: records the thread's last error

namespace {

// Orientation code of a vertex-sharing pair (t, s) from the 3 x 3 vertex match matrix m[i][j] (i: test corner,
// j: trial corner): 100000 coincident; 200000 + 1000 ia + 100 ib + 10 ja + jb for an edge, the two matches taken
// in (i, j) scan order and ordered by trial corner; 300000 + 1000 ia + 10 ja for a shared vertex.
int pair_code(const int* ft, const int* fs) {
  int n = 0, first = -1, second = -1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (ft[i] == fs[j]) {
        if (first < 0)
          first = 3 * i + j;
        else if (second < 0)
          second = 3 * i + j;
        ++n;
      }
  if (n >= 3) return 100000;
  int ia = first / 3, ja = first % 3;
  if (n == 2) {
    int ib = second / 3, jb = second % 3;
    if (jb < ja) {
      std::swap(ia, ib);
      std::swap(ja, jb);
    }
    return 200000 + ia * 1000 + ib * 100 + ja * 10 + jb;
  }
  return 300000 + ia * 1000 + ja * 10;
}

}  // namespace

extern "C" {

int bem_host_vertex_pairs(int F, int V, const int* faces, long long capacity, int* test, int* trial, int* rule,
                          int* codes, int* n_codes, long long* n_pairs) {
  if (F <= 0 || V <= 0 || !faces || !test || !trial || !rule || !codes || !n_codes || !n_pairs)
    return bem_fail_message(BEM_E_ARG, "bem_host_vertex_pairs: invalid arguments");
  try {
    for (long long i = 0; i < 3LL * F; ++i)
      if (faces[i] < 0 || faces[i] >= V) return bem_fail_message(BEM_E_ARG, "face vertex id out of range");
    // vertex -> incident faces (CSR, faces ascending)
    std::vector<int> ptr(V + 1, 0), inc(3 * (size_t)F);
    for (long long i = 0; i < 3LL * F; ++i) ++ptr[faces[i] + 1];
    for (int v = 0; v < V; ++v) ptr[v + 1] += ptr[v];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int t = 0; t < F; ++t)
      for (int c = 0; c < 3; ++c) inc[fill[faces[3 * t + c]]++] = t;
    // pairs grouped by test element, trial ascending; codes mapped to their rank among the distinct codes
    std::vector<int> cand;
    long long n = 0;
    std::vector<int> code_of;  // per pair, before ranking
    code_of.reserve((size_t)std::min<long long>(capacity, 1LL << 26));
    for (int t = 0; t < F; ++t) {
      cand.clear();
      for (int c = 0; c < 3; ++c) {
        const int v = faces[3 * t + c];
        cand.insert(cand.end(), inc.begin() + ptr[v], inc.begin() + ptr[v + 1]);
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      if (n + (long long)cand.size() > capacity) return bem_fail_message(BEM_E_ARG, "pair capacity too small");
      for (int s : cand) {
        test[n] = t;
        trial[n] = s;
        code_of.push_back(pair_code(faces + 3 * t, faces + 3 * s));
        ++n;
      }
    }
    // rank of each code among the distinct codes, without sorting the pairs' codes: the codes map one to one and in
    // order onto 91 slots (coincident; edge by (ia, ib, ja, jb); vertex by (ia, ja)), marked and prefix-counted
    auto slot_of = [](int code) {
      if (code == 100000) return 0;
      if (code < 300000) {
        const int d = code - 200000;
        return 1 + ((d / 1000 * 3 + d / 100 % 10) * 3 + d / 10 % 10) * 3 + d % 10;
      }
      const int d = code - 300000;
      return 82 + d / 1000 * 3 + d / 10 % 10;
    };
    int slot_code[91] = {0}, rank[91];
    bool present[91] = {false};
    for (long long q = 0; q < n; ++q) {
      const int sl = slot_of(code_of[q]);
      present[sl] = true;
      slot_code[sl] = code_of[q];
    }
    int nu = 0;
    for (int sl = 0; sl < 91; ++sl)
      if (present[sl]) {
        if (nu == 64) return bem_fail_message(BEM_E_ARG, "more than 64 orientation codes");
        codes[nu] = slot_code[sl];
        rank[sl] = nu++;
      }
    *n_codes = nu;
    for (long long q = 0; q < n; ++q) rule[q] = rank[slot_of(code_of[q])];
    *n_pairs = n;
  } catch (const std::exception& ex) {
    return bem_fail_message(BEM_E_NOMEM, std::string("bem_host_vertex_pairs: ") + ex.what());
  }
  return BEM_OK;
}

int bem_host_slot_tables(int F, int V, const int* faces, const int* order, int pend_slots, int* slot_count,
                         uint16_t* slot_start, uint16_t* slot_inc, int* slot_vertex, int* slot_pend, int* slot_prev,
                         int* elem_pos, int* chunk_start, int* vertex_last, int* vstar_ptr, int* vstar_row) {
  if (F <= 0 || V <= 0 || !faces || !order || pend_slots < 0 || pend_slots > BEM_PEND_SLOTS || !slot_count ||
      !slot_start || !slot_inc || !slot_vertex || !slot_pend || !slot_prev || !elem_pos || !chunk_start ||
      !vertex_last || !vstar_ptr || !vstar_row)
    return bem_fail_message(BEM_E_ARG, "bem_host_slot_tables: invalid arguments");
  try {
    const int n_chunks = (F + 63) / 64;
    for (long long i = 0; i < 3LL * F; ++i)
      if (faces[i] < 0 || faces[i] >= V) return bem_fail_message(BEM_E_ARG, "face vertex id out of range");
    for (long long i = 0; i < 64LL * n_chunks; ++i)
      if (i < F ? (order[i] < 0 || order[i] >= F) : order[i] != -1)
        return bem_fail_message(BEM_E_ARG, "order must be a permutation of the faces padded with -1");
    std::memset(slot_count, 0, sizeof(int) * n_chunks);
    std::memset(slot_start, 0, sizeof(uint16_t) * (size_t)n_chunks * BEM_SLOT_STRIDE);
    std::memset(slot_inc, 0, sizeof(uint16_t) * (size_t)n_chunks * BEM_SLOT_INC);
    std::memset(slot_vertex, 0, sizeof(int) * (size_t)n_chunks * BEM_SLOT_INC);
    for (size_t i = 0; i < (size_t)n_chunks * BEM_SLOT_INC; ++i) {
      slot_pend[i] = BEM_PEND_NONE | BEM_PEND_TO_STRIP;
      slot_prev[i] = -1;
    }
    // per chunk: its distinct vertices ascending, and the incidence codes (lane << 2 | corner) sorted by vertex,
    // then code
    std::vector<std::vector<int>> chunk_verts(n_chunks);
    std::vector<std::pair<int, int>> vc;  // (vertex, code)
    for (int c = 0; c < n_chunks; ++c) {
      vc.clear();
      for (int lane = 0; lane < 64; ++lane) {
        const int e = order[64 * c + lane];
        if (e < 0) continue;
        for (int k = 0; k < 3; ++k) vc.emplace_back(faces[3 * e + k], (lane << 2) | k);
      }
      std::sort(vc.begin(), vc.end());
      std::vector<int>& uv = chunk_verts[c];
      uint16_t* st = slot_start + (size_t)c * BEM_SLOT_STRIDE;
      uint16_t* si = slot_inc + (size_t)c * BEM_SLOT_INC;
      for (size_t i = 0; i < vc.size(); ++i) {
        si[i] = (uint16_t)vc[i].second;
        if (uv.empty() || uv.back() != vc[i].first) uv.push_back(vc[i].first);
        st[uv.size()] = (uint16_t)(i + 1);
      }
      slot_count[c] = (int)uv.size();
    }
    std::vector<long long> first(V, -1), last(V, -1);
    for (int c = 0; c < n_chunks; ++c)
      for (int v : chunk_verts[c]) {
        if (first[v] < 0) first[v] = c;
        last[v] = c;
      }
    // LDS slots for vertices used by several chunks: scan the live ranges [first, last] by start (then end, then
    // vertex), keep as many as fit pend_slots at every chunk, dropping the covering range that ends last (ties:
    // the lowest vertex id) when over capacity
    std::vector<int> multi;
    for (int v = 0; v < V; ++v)
      if (first[v] >= 0 && last[v] > first[v]) multi.push_back(v);
    std::stable_sort(multi.begin(), multi.end(), [&](int a, int b) {
      return first[a] != first[b] ? first[a] < first[b] : last[a] < last[b];
    });
    std::vector<char> keep(V, 0);
    auto worse = [&](const std::pair<long long, int>& a, const std::pair<long long, int>& b) {
      // max-heap on last chunk; on ties the lower vertex id is popped first
      return a.first != b.first ? a.first < b.first : a.second > b.second;
    };
    std::priority_queue<std::pair<long long, int>, std::vector<std::pair<long long, int>>, decltype(worse)> active(
        worse);
    std::vector<long long> ends(n_chunks + 1, 0);
    long long live = 0, scan = 0;
    for (int v : multi) {
      const long long fv = first[v], lv = last[v];
      while (scan < fv) live -= ends[scan++];
      keep[v] = 1;
      active.emplace(lv, v);
      ++live;
      ++ends[lv];
      if (live > pend_slots) {
        while (!keep[active.top().second] || last[active.top().second] < fv) active.pop();
        const int u = active.top().second;
        active.pop();
        keep[u] = 0;
        --live;
        --ends[last[u]];
      }
    }
    std::vector<long long> pend(V, -1), last_seen(V, -1);
    std::vector<int> freelist;
    for (int i = pend_slots - 1; i >= 0; --i) freelist.push_back(i);
    std::vector<std::vector<int>> release(n_chunks);
    for (int c = 0; c < n_chunks; ++c) {
      const std::vector<int>& uv = chunk_verts[c];
      for (int v : uv)
        if (first[v] == c && keep[v]) {
          pend[v] = freelist.back();
          freelist.pop_back();
          release[last[v]].push_back((int)pend[v]);
        }
      for (size_t k = 0; k < uv.size(); ++k) {
        const int v = uv[k];
        const bool seen = first[v] < c, in_lds = pend[v] >= 0;
        int code = in_lds ? (int)pend[v] : BEM_PEND_NONE;
        if (in_lds && seen) code |= BEM_PEND_FROM_LDS;
        if (!in_lds || last[v] == c) code |= BEM_PEND_TO_STRIP;
        const size_t o = (size_t)c * BEM_SLOT_INC + k;
        slot_pend[o] = code;
        slot_vertex[o] = 2 * v + ((seen && !in_lds) ? 1 : 0);
        slot_prev[o] = (int)last_seen[v];
        last_seen[v] = c;
      }
      freelist.insert(freelist.end(), release[c].begin(), release[c].end());
    }
    // element positions, the first chunk a test element's wave visits, each vertex's last chunk
    for (int i = 0; i < F; ++i) elem_pos[order[i]] = i;
    std::vector<int> vmin(V, n_chunks);
    for (int t = 0; t < F; ++t)
      for (int k = 0; k < 3; ++k) vmin[faces[3 * t + k]] = std::min(vmin[faces[3 * t + k]], elem_pos[t] / 64);
    for (int t = 0; t < F; ++t) {
      int m = elem_pos[t] / 64;
      for (int k = 0; k < 3; ++k) m = std::min(m, vmin[faces[3 * t + k]]);
      chunk_start[t] = m;
    }
    for (int v = 0; v < V; ++v) vertex_last[v] = (int)last[v];
    // vertex -> strip rows 3t + k, rows ascending
    std::memset(vstar_ptr, 0, sizeof(int) * ((size_t)V + 1));
    for (long long r = 0; r < 3LL * F; ++r) ++vstar_ptr[faces[r] + 1];
    for (int v = 0; v < V; ++v) vstar_ptr[v + 1] += vstar_ptr[v];
    std::vector<int> pos(vstar_ptr, vstar_ptr + V);
    for (long long r = 0; r < 3LL * F; ++r) vstar_row[pos[faces[r]]++] = (int)r;
  } catch (const std::exception& ex) {
    return bem_fail_message(BEM_E_NOMEM, std::string("bem_host_slot_tables: ") + ex.what());
  }
  return BEM_OK;
}

// Float32 rows as JSON text. A value is written with nine significant digits, which is enough for every float32
// value to read back exactly even when the last digit is off by one (the nine-digit quantum is at most 1e-8 of
// the value, under half the float32 spacing of 6e-8): the digits come from one double product with a power of
// ten, rounded to the nearest integer, and are printed in %g style (exponent form below 1e-4 and from 1e9).
namespace {
double pow10_table(int k) {  // 10^k for |k| <= 60, from exact powers (10^22 is the largest exact double)
  static const double kPos[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                  1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  double r = 1.0;
  int a = k < 0 ? -k : k;
  while (a > 22) {
    r *= 1e22;
    a -= 22;
  }
  r *= kPos[a];
  return k < 0 ? 1.0 / r : r;
}

char* put_float32(char* o, float x) {
  if (x == 0.0f) {
    if (std::signbit(x)) *o++ = '-';
    *o++ = '0';
    return o;
  }
  double a = (double)x;
  if (a < 0) {
    *o++ = '-';
    a = -a;
  }
  int e = (int)std::floor(std::log10(a));
  long long m = std::llround(a * pow10_table(8 - e));
  if (m >= 1000000000LL) {
    ++e;
    m = std::llround(a * pow10_table(8 - e));
  } else if (m < 100000000LL) {
    --e;
    m = std::llround(a * pow10_table(8 - e));
  }
  if (m >= 1000000000LL) {  // rounded up to the next decade
    m /= 10;
    ++e;
  }
  char d[9];
  for (int i = 8; i >= 0; --i) {
    d[i] = (char)('0' + m % 10);
    m /= 10;
  }
  int nd = 9;
  while (nd > 1 && d[nd - 1] == '0') --nd;
  if (e < -4 || e >= 9) {
    *o++ = d[0];
    if (nd > 1) {
      *o++ = '.';
      for (int i = 1; i < nd; ++i) *o++ = d[i];
    }
    *o++ = 'e';
    *o++ = e < 0 ? '-' : '+';
    const int ae = e < 0 ? -e : e;
    if (ae >= 100) *o++ = (char)('0' + ae / 100);
    *o++ = (char)('0' + (ae / 10) % 10);
    *o++ = (char)('0' + ae % 10);
  } else if (e >= 0) {
    for (int i = 0; i <= e; ++i) *o++ = i < nd ? d[i] : '0';
    if (nd > e + 1) {
      *o++ = '.';
      for (int i = e + 1; i < nd; ++i) *o++ = d[i];
    }
  } else {
    *o++ = '0';
    *o++ = '.';
    for (int i = 0; i < -e - 1; ++i) *o++ = '0';
    for (int i = 0; i < nd; ++i) *o++ = d[i];
  }
  return o;
}
}  // namespace

long long bem_host_format_f32_json(const float* v, long long n, char* out, long long cap) {
  if (n < 0 || (n > 0 && !v) || !out) return -1;
  if (cap < 16 * n + 2) return -1;
  char* o = out;
  *o++ = '[';
  for (long long i = 0; i < n; ++i) {
    if (!std::isfinite(v[i])) return -2;
    if (i) *o++ = ',';
    o = put_float32(o, v[i]);
  }
  *o++ = ']';
  return (long long)(o - out);
}

}  // extern "C"
