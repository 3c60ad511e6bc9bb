// Atom data files of the reference's enrichment format and their numerical
// radial functions.
//
// Reads the XML atom file of AtomSphericalDataEnrichment (src/atoms/
// AtomSphericalDataEnrichment.cpp:511-640): root <atom> in the namespace
// http://www.dft-efe.com/dft-efe; single metadata elements (symbol, Z, charge,
// NR); the radial grid <r> of NR points; per element of a field (e.g. every
// <orbital>) the children <qNumbers> (n l m), <values> (NR radial values) and
// <cutoffInfo> (cutoff, smoothness), with the checks of its
// processSphericalDataFromXMLNodeData (:347-473).  Each radial function is
// interpolated by the C^2 cubic spline with zero end slopes (the defaults of
// utils::Spline, src/utils/Spline.h:87-95; evaluation and extrapolation as
// Spline.cpp:490-518) and evaluated as SphericalDataNumerical does
// (src/atoms/SphericalDataNumerical.cpp:40-140): R(r) Y_lm(theta, phi) f_c(r),
// zero past cutoff (1 + 1 / smoothness), with the smooth cutoff of
// SmoothCutoffFunctions.cpp:25-70 and the real spherical harmonics of
// SphericalHarmonicFunctions.cpp:348-430 (no Condon-Shortley phase).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>
#include <type_traits>

#include "api_util.hpp"
#include "atom.hpp"

using namespace dfx;

namespace
{
  const char *kNamespace = "http://www.dft-efe.com/dft-efe";

  // ------------------------------------------------------- minimal XML
  // Elements, attributes, character data, comments, CDATA, processing
  // instructions and the five predefined entities; names keep their prefix.
  struct XNode
  {
    std::string                                      name; // qualified
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string                                      text;
    std::vector<XNode>                               kids;
  };

  struct XParser
  {
    const std::string &s;
    const std::string &file;
    size_t             i = 0;

    [[noreturn]] void fail(const std::string &what) const
    {
      throw Error(DFX_E_ARG, "Error parsing the XML file:" + file + " (" + what + " at byte " +
                               std::to_string(i) + ")");
    }
    bool starts(const char *t) const { return s.compare(i, strlen(t), t) == 0; }
    void skipTo(const char *t)
    {
      const size_t j = s.find(t, i);
      if (j == std::string::npos)
        fail(std::string("missing ") + t);
      i = j + strlen(t);
    }
    void ws()
    {
      while (i < s.size() && isspace((unsigned char)s[i]))
        ++i;
    }
    std::string name()
    {
      const size_t j = i;
      while (i < s.size() && !isspace((unsigned char)s[i]) && s[i] != '>' && s[i] != '/' &&
             s[i] != '=')
        ++i;
      if (i == j)
        fail("empty name");
      return s.substr(j, i - j);
    }
    static std::string decode(const std::string &t)
    {
      std::string o;
      for (size_t k = 0; k < t.size(); ++k)
        {
          if (t[k] != '&')
            {
              o += t[k];
              continue;
            }
          static const std::pair<const char *, char> ents[] = {
            {"&lt;", '<'}, {"&gt;", '>'}, {"&amp;", '&'}, {"&quot;", '"'}, {"&apos;", '\''}};
          bool hit = false;
          for (const auto &e : ents)
            if (t.compare(k, strlen(e.first), e.first) == 0)
              {
                o += e.second;
                k += strlen(e.first) - 1;
                hit = true;
                break;
              }
          if (!hit)
            o += '&';
        }
      return o;
    }
    void misc() // comments, processing instructions, doctype, whitespace
    {
      for (;;)
        {
          ws();
          if (starts("<?"))
            skipTo("?>");
          else if (starts("<!--"))
            skipTo("-->");
          else if (starts("<!DOCTYPE"))
            skipTo(">");
          else
            return;
        }
    }
    int depth = 0;
    XNode element()
    {
      if (!starts("<"))
        fail("expected an element");
      if (++depth > 64)
        fail("elements nested more than 64 deep");
      ++i;
      XNode n;
      n.name = name();
      for (;;)
        {
          ws();
          if (starts("/>"))
            {
              i += 2;
              --depth;
              return n;
            }
          if (starts(">"))
            {
              ++i;
              break;
            }
          std::string an = name();
          ws();
          if (!starts("="))
            fail("attribute without value");
          ++i;
          ws();
          const char qc = i < s.size() ? s[i] : 0;
          if (qc != '"' && qc != '\'')
            fail("unquoted attribute");
          const size_t e = s.find(qc, i + 1);
          if (e == std::string::npos)
            fail("unterminated attribute");
          n.attrs.emplace_back(an, decode(s.substr(i + 1, e - i - 1)));
          i = e + 1;
        }
      for (;;)
        {
          if (i >= s.size())
            fail("unterminated element <" + n.name + ">");
          if (starts("</"))
            {
              i += 2;
              const std::string cn = name();
              if (cn != n.name)
                fail("</" + cn + "> closes <" + n.name + ">");
              ws();
              if (!starts(">"))
                fail("malformed end tag");
              ++i;
              --depth;
              return n;
            }
          if (starts("<!--"))
            skipTo("-->");
          else if (starts("<![CDATA["))
            {
              i += 9;
              const size_t e = s.find("]]>", i);
              if (e == std::string::npos)
                fail("unterminated CDATA");
              n.text += s.substr(i, e - i);
              i = e + 3;
            }
          else if (starts("<?"))
            skipTo("?>");
          else if (starts("<"))
            n.kids.push_back(element());
          else
            {
              const size_t e = s.find('<', i);
              n.text += decode(s.substr(i, (e == std::string::npos ? s.size() : e) - i));
              i = e == std::string::npos ? s.size() : e;
            }
        }
    }
  };

  std::string localName(const std::string &q)
  {
    const size_t c = q.find(':');
    return c == std::string::npos ? q : q.substr(c + 1);
  }
  std::string prefixOf(const std::string &q)
  {
    const size_t c = q.find(':');
    return c == std::string::npos ? std::string() : q.substr(0, c);
  }
  // namespace URI of element n given its ancestors' declarations
  std::string nsOf(const XNode &n, std::map<std::string, std::string> scope)
  {
    for (const auto &a : n.attrs)
      {
        if (a.first == "xmlns")
          scope[""] = a.second;
        else if (a.first.rfind("xmlns:", 0) == 0)
          scope[a.first.substr(6)] = a.second;
      }
    const auto it = scope.find(prefixOf(n.name));
    return it == scope.end() ? std::string() : it->second;
  }
  std::map<std::string, std::string> scopeOf(const XNode &n,
                                             std::map<std::string, std::string> scope)
  {
    for (const auto &a : n.attrs)
      {
        if (a.first == "xmlns")
          scope[""] = a.second;
        else if (a.first.rfind("xmlns:", 0) == 0)
          scope[a.first.substr(6)] = a.second;
      }
    return scope;
  }

  std::string trim(const std::string &t)
  {
    size_t a = 0, b = t.size();
    while (a < b && isspace((unsigned char)t[a]))
      ++a;
    while (b > a && isspace((unsigned char)t[b - 1]))
      --b;
    return t.substr(a, b - a);
  }

  // whitespace-separated doubles; false if a token is not a whole number
  bool toDoubles(const std::string &t, std::vector<double> &out)
  {
    out.clear();
    std::istringstream is(t);
    std::string        tok;
    while (is >> tok)
      {
        char *end = nullptr;
        errno     = 0;
        const double v = strtod(tok.c_str(), &end);
        if (end != tok.c_str() + tok.size() || errno == ERANGE)
          return false;
        out.push_back(v);
      }
    return true;
  }
  bool toInts(const std::string &t, std::vector<int> &out)
  {
    out.clear();
    std::istringstream is(t);
    std::string        tok;
    while (is >> tok)
      {
        char *end = nullptr;
        errno     = 0;
        const long v = strtol(tok.c_str(), &end, 10);
        if (end != tok.c_str() + tok.size() || errno == ERANGE || v < INT32_MIN || v > INT32_MAX)
          return false;
        out.push_back((int)v);
      }
    return true;
  }

  // C^2 cubic spline through (x_i, y_i) with a first- or second-derivative
  // condition at each end (utils::Spline's bd_type; default: slopes 0):
  // c_i = f''(x_i) / 2 from the tridiagonal system (Thomas algorithm; the
  // matrix is strictly diagonally dominant), then b_i and d_i per interval.
  void cubicSpline(const std::vector<double> &x, dfx_atom::Function &F, bool leftSecond = false,
                   double leftValue = 0.0, bool rightSecond = false, double rightValue = 0.0)
  {
    const size_t n = x.size();
    const auto  &y = F.y;
    F.b.assign(n, 0.0);
    F.c.assign(n, 0.0);
    F.d.assign(n, 0.0);
    std::vector<double> lo(n, 0.0), di(n, 0.0), up(n, 0.0), rhs(n, 0.0);
    for (size_t i = 0; i + 1 < n; ++i)
      {
        const double h  = x[i + 1] - x[i];
        const double dy = (y[i + 1] - y[i]) / h;
        // row i gets 2h c_i + h c_{i+1} += 3 dy ; row i+1 gets h c_i + 2h c_{i+1} -= 3 dy
        di[i] += 2.0 * h;
        up[i] += h;
        rhs[i] += 3.0 * dy;
        lo[i + 1] += h;
        di[i + 1] += 2.0 * h;
        rhs[i + 1] -= 3.0 * dy;
      }
    if (leftSecond) // f''(x_0) given: c_0 = value / 2
      {
        di[0]  = 1.0;
        up[0]  = 0.0;
        rhs[0] = 0.5 * leftValue;
      }
    else // f'(x_0) given
      rhs[0] -= 3.0 * leftValue;
    if (rightSecond)
      {
        lo[n - 1]  = 0.0;
        di[n - 1]  = 1.0;
        rhs[n - 1] = 0.5 * rightValue;
      }
    else
      rhs[n - 1] += 3.0 * rightValue;
    for (size_t i = 1; i < n; ++i)
      {
        const double w = lo[i] / di[i - 1];
        di[i] -= w * up[i - 1];
        rhs[i] -= w * rhs[i - 1];
      }
    F.c[n - 1] = rhs[n - 1] / di[n - 1];
    for (size_t i = n - 1; i-- > 0;)
      F.c[i] = (rhs[i] - up[i] * F.c[i + 1]) / di[i];
    for (size_t i = 0; i + 1 < n; ++i)
      {
        const double h = x[i + 1] - x[i];
        F.d[i]         = (F.c[i + 1] - F.c[i]) / (3.0 * h);
        F.b[i]         = (y[i + 1] - y[i]) / h - h * (2.0 * F.c[i] + F.c[i + 1]) / 3.0;
      }
    // slope at the last point from the last interval's cubic; no cubic term
    // beyond it (right extrapolation is quadratic)
    const double h = x[n - 1] - x[n - 2];
    F.b[n - 1]     = (3.0 * F.d[n - 2] * h + 2.0 * F.c[n - 2]) * h + F.b[n - 2];
    F.d[n - 1]     = 0.0;
    // left extrapolation: linear after a slope condition, else quadratic
    F.cLeft = leftSecond ? F.c[0] : 0.0;
  }

  // f(x) = R(r) Y_lm(x / r) f_c(r) and its gradient at every point (rows of
  // P, 3 values each) relative to the origin; l <= 2
  __global__ void atomEvalK(SplineDev S, int l, int m, double rc, double sm, int mixed,
                            double tail, double ox, double oy, double oz, const double *P,
                            u32 nPts, double *val, double *grad)
  {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPts)
      return;
    const double dx[3] = {P[3 * (u64)i] - ox, P[3 * (u64)i + 1] - oy, P[3 * (u64)i + 2] - oz};
    double       f, g[3];
    if (mixed)
      {
        // SphericalDataMixed (src/atoms/SphericalDataMixed.cpp:43-69): the
        // spline up to the last knot, tail / r beyond, times Y_00
        const double y00 = 0.28209479177387814;
        const double r   = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double       R, dR;
        if (r <= S.x[S.n - 1])
          splineEval(S, r, R, dR);
        else
          {
            R  = tail / r;
            dR = -tail / (r * r);
          }
        f = y00 * R;
        for (int k = 0; k < 3; ++k)
          g[k] = r > 0.0 ? y00 * dR * dx[k] / r : 0.0;
      }
    else
      radialYlm(S, l, m, rc, sm, dx, f, g);
    val[i] = f;
    if (grad)
      {
        grad[3 * (u64)i]     = g[0];
        grad[3 * (u64)i + 1] = g[1];
        grad[3 * (u64)i + 2] = g[2];
      }
  }

  const dfx_atom::Function &functionOf(const dfx_atom *a, const char *field, uint32_t k)
  {
    DFX_REQUIRE(a && field, "null atom data or field name");
    const auto it = a->fields.find(field);
    DFX_REQUIRE(it != a->fields.end(), std::string("field ") + field + " was not read from " + a->file);
    DFX_REQUIRE(k < it->second.size(), "function index out of range");
    return it->second[k];
  }
} // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int dfx_atom_read(const char *path, const char *const *fieldNames, uint32_t nFields,
                  dfx_atom **out)
{
  return guard(nullptr, [&] {
    DFX_REQUIRE(path && out && (nFields == 0 || fieldNames), "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f)
      throw Error(DFX_E_ARG, std::string("Error parsing the XML file:") + path +
                               " (cannot open)");
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const std::string file(path);
    XParser           xp{text, file};
    xp.misc();
    XNode root = xp.element();
    xp.misc();
    if (xp.i != text.size())
      xp.fail("content after the root element");
    auto A  = std::make_unique<dfx_atom>();
    A->file = file;
    const auto scope = scopeOf(root, {});
    if (localName(root.name) != "atom" || nsOf(root, {}) != kNamespace)
      throw Error(DFX_E_ARG, "XPath expression /dft-efe:atom not found in " + file);
    // children of <atom> in the dft-efe namespace, by local name
    std::map<std::string, std::vector<const XNode *>> kids;
    for (const XNode &c : root.kids)
      if (nsOf(c, scope) == kNamespace)
        {
          kids[localName(c.name)].push_back(&c);
          A->text[localName(c.name)].push_back(c.text);
        }
    auto single = [&](const char *name) -> std::string {
      const auto it = kids.find(name);
      if (it == kids.end())
        throw Error(DFX_E_ARG, std::string("XPath expression /dft-efe:atom/dft-efe:") + name +
                                 " not found.");
      if (it->second.size() != 1)
        throw Error(DFX_E_ARG, std::string("Found more than one /dft-efe:atom/dft-efe:") + name +
                                 " element in " + file);
      return it->second[0]->text;
    };
    // number of radial points and the grid
    std::vector<int> nr;
    if (!toInts(trim(single("NR")), nr) || nr.size() != 1)
      throw Error(DFX_E_ARG, "Error while converting /dft-efe:atom/dft-efe:NR element in " + file +
                               " to int");
    if (nr[0] <= 0)
      throw Error(DFX_E_ARG, "Non-positive integer found for/dft-efe:atom/dft-efe:NR element in " +
                               file);
    const size_t NR = (size_t)nr[0];
    if (!toDoubles(single("r"), A->r))
      throw Error(DFX_E_ARG, "Error while converting /dft-efe:atom/dft-efe:r element in " + file +
                               " to double");
    if (A->r.size() != NR)
      throw Error(DFX_E_ARG, "Mismatch in number of radial points specified and the number of "
                             " radial points provided in " + file);
    // the spline needs a strictly increasing grid of at least two points
    DFX_REQUIRE(NR >= 2, "the radial grid needs at least two points: " + file);
    for (size_t i = 0; i + 1 < NR; ++i)
      DFX_REQUIRE(A->r[i + 1] > A->r[i], "the radial grid is not strictly increasing: " + file);
    for (uint32_t fi = 0; fi < nFields; ++fi)
      {
        DFX_REQUIRE(fieldNames[fi], "null field name");
        const std::string fname = fieldNames[fi];
        const std::string xpath = "/dft-efe:atom/dft-efe:" + fname;
        const auto        it    = kids.find(fname);
        if (it == kids.end())
          throw Error(DFX_E_ARG, "XPath expression " + xpath + "/dft-efe:values not found.");
        std::vector<dfx_atom::Function> fun;
        std::set<std::vector<int>>      qset;
        for (const XNode *e : it->second)
          {
            const auto escope = scopeOf(*e, scope);
            std::map<std::string, std::vector<const XNode *>> sub;
            for (const XNode &c : e->kids)
              if (nsOf(c, escope) == kNamespace)
                sub[localName(c.name)].push_back(&c);
            for (const char *t : {"values", "qNumbers", "cutoffInfo"})
              if (sub[t].size() != 1)
                throw Error(DFX_E_ARG, std::string("Mismatch in number of \"values\", \"qNumbers\" "
                                                   "and \"cutoffInfo\" child elements found for "
                                                   "element ") +
                                         xpath + " in file " + file);
            dfx_atom::Function F;
            if (!toDoubles(sub["values"][0]->text, F.y))
              throw Error(DFX_E_ARG, "Error while converting values in " + xpath + " element in " +
                                       file + " to double");
            if (F.y.size() != NR)
              throw Error(DFX_E_ARG, "Mismatch in number of points specified and number of points "
                                     "provided in " + xpath + " element in " + file);
            std::vector<int> q;
            if (!toInts(sub["qNumbers"][0]->text, q))
              throw Error(DFX_E_ARG, "Error while converting quantum numbers in " + xpath +
                                       " element in " + file + " to double");
            if (q.size() != 3)
              throw Error(DFX_E_ARG, "Expected three quantum number (i.e., n, l, m) in "
                                     "\"qNumbers\" child element in " + xpath + " element in " +
                                       file);
            if (!(q[2] >= -q[1] && q[2] <= q[1]))
              throw Error(DFX_E_ARG, "Magnetic quantum number (m) found outside of -l and +l "
                                     "(l = angular quantum number) in " + xpath + " element in " +
                                       file);
            qset.insert(q);
            std::vector<double> co;
            if (!toDoubles(sub["cutoffInfo"][0]->text, co))
              throw Error(DFX_E_ARG, "Error while converting cutoff info in " + xpath +
                                       " element in " + file + " to double");
            if (co.size() != 2)
              throw Error(DFX_E_ARG, "Expected two values (cutoff and smoothness factor) in  "
                                     "\"cutoffInfo\" child element in " + xpath + " element in " +
                                       file);
            F.q[0] = q[0], F.q[1] = q[1], F.q[2] = q[2];
            F.cutoff     = co[0];
            F.smoothness = co[1];
            cubicSpline(A->r, F);
            fun.push_back(std::move(F));
          }
        if (qset.size() != fun.size())
          throw Error(DFX_E_ARG, "Found repeated quantum numbers while processing " + xpath +
                                   " element in " + file);
        // in (n, l, m) order, every (n, l) lists m = -l .. l
        std::vector<std::vector<int>> qs(qset.begin(), qset.end());
        for (size_t i = 0; i < qs.size();)
          {
            const int l = qs[i][1];
            for (int m = -l; m <= l; ++m, ++i)
              if (i >= qs.size() || qs[i][2] != m)
                throw Error(DFX_E_ARG, "The quantum number m for " + xpath + " in the xml file " +
                                         file + " are not in correct order (ascending m).");
          }
        A->fields[fname] = std::move(fun);
      }
    *out = A.release();
    return 0;
  });
}

// Pseudopotential files, UPF v2 (AtomSphericalDataPSP, src/atoms/
// AtomSphericalDataPSP.cpp:291-830): PP_HEADER attributes as metadata (with
// its checks of mesh_size, z_valence, l_max, number_of_proj), the mesh
// PP_MESH/PP_R, and per field (Defaults.cpp:41-47: vlocal -> PP_LOCAL, beta
// -> PP_NONLOCAL/PP_BETA.*, pswfc -> PP_PSWFC/PP_CHI.*, nlcc -> PP_NLCC,
// rhoatom -> PP_RHOATOM) the radial data with the reference's transforms:
// Ry -> Ha for vlocal and D_ij, beta and pswfc divided by r, rhoatom by
// 4 pi r^2 (first value copied from the second), everything but beta and
// pswfc divided by Y_00; support from the last |value| >= 1e-10 (rhoatom: the
// whole mesh), smoothness 1e10; vlocal is mixed: the spline of the points
// within 10.0001 bohr (f'' at the first point = the first divided
// difference, f' at the last = -v / r) and -|z_valence| / (Y_00 r) beyond.
int dfx_atom_read_upf(const char *path, const char *const *fieldNames, uint32_t nFields,
                      dfx_atom **out)
{
  return guard(nullptr, [&] {
    DFX_REQUIRE(path && out && (nFields == 0 || fieldNames), "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f)
      throw Error(DFX_E_ARG, std::string("Error parsing the XML file:") + path +
                               " (cannot open)");
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const std::string file(path);
    XParser           xp{text, file};
    xp.misc();
    XNode root = xp.element();
    xp.misc();
    if (xp.i != text.size())
      xp.fail("content after the root element");
    if (localName(root.name) != "UPF")
      throw Error(DFX_E_ARG, "XPath expression /UPF not found in " + file);
    auto A  = std::make_unique<dfx_atom>();
    A->file = file;
    auto children = [&](const XNode &n, const char *name) {
      std::vector<const XNode *> v;
      for (const XNode &c : n.kids)
        if (c.name == name)
          v.push_back(&c);
      return v;
    };
    auto single = [&](const XNode &n, const char *name, const std::string &xpath) {
      const auto v = children(n, name);
      if (v.empty())
        throw Error(DFX_E_ARG, "XPath expression " + xpath + " not found.");
      if (v.size() != 1)
        throw Error(DFX_E_ARG, "Found more than one " + xpath + " element in " + file);
      return v[0];
    };
    // header
    const XNode *hdr = single(root, "PP_HEADER", "/UPF/PP_HEADER");
    int    meshSize = 0, lmax = 0, nProj = 0;
    double zv       = 0.0;
    bool   coreCorr = false;
    auto   num = [&](const std::string &v, const char *what, auto &dst) {
      std::vector<double> d;
      if (!toDoubles(v, d) || d.size() != 1)
        throw Error(DFX_E_ARG, std::string("Error while converting /UPF/PP_HEADER element in ") +
                                 file + " (" + what + ")");
      using D = std::remove_reference_t<decltype(dst)>;
      if (std::is_integral<D>::value && !(std::fabs(d[0]) < 2147483647.0))
        throw Error(DFX_E_ARG, std::string("Error while converting /UPF/PP_HEADER element in ") +
                                 file + " to int (" + what + ")");
      dst = (D)d[0];
      if ((double)dst != d[0])
        throw Error(DFX_E_ARG, std::string("Error while converting /UPF/PP_HEADER element in ") +
                                 file + " to int (" + what + ")");
    };
    for (const auto &at : hdr->attrs)
      {
        const std::string v = trim(at.second);
        A->text[at.first].push_back(v);
        if (at.first == "mesh_size")
          {
            num(v, "mesh_size", meshSize);
            if (meshSize <= 0)
              throw Error(DFX_E_ARG, "Non-positive numRadial points found for/UPF/PP_HEADER "
                                     "element in " + file);
          }
        else if (at.first == "z_valence")
          {
            num(v, "z_valence", zv);
            if (!(zv > 0))
              throw Error(DFX_E_ARG, "Non-positive z valance found for/UPF/PP_HEADER element in " +
                                       file);
          }
        else if (at.first == "l_max")
          {
            num(v, "l_max", lmax);
            if (lmax < 0)
              throw Error(DFX_E_ARG, "Non-positive lmax found for/UPF/PP_HEADER element in " + file);
          }
        else if (at.first == "number_of_proj")
          {
            num(v, "number_of_proj", nProj);
            if (nProj < 0)
              throw Error(DFX_E_ARG, "Non-positive lmax found for/UPF/PP_HEADER element in " + file);
          }
        else if (at.first == "core_correction")
          {
            std::string b;
            for (char ch : v)
              b += (char)tolower((unsigned char)ch);
            if (b == "t" || b == "true" || b == ".true." || b == "1")
              coreCorr = true;
            else if (b == "f" || b == "false" || b == ".false." || b == "0")
              coreCorr = false;
            else
              throw Error(DFX_E_ARG, "Error while converting core_correction in " + file);
          }
      }
    DFX_REQUIRE(meshSize > 0, "no mesh_size in /UPF/PP_HEADER of " + file);
    // mesh
    const XNode *mesh = single(root, "PP_MESH", "/UPF/PP_MESH");
    const XNode *pr   = single(*mesh, "PP_R", "/UPF/PP_MESH/PP_R");
    if (!toDoubles(pr->text, A->r))
      throw Error(DFX_E_ARG, "Error while converting /UPF/PP_MESH/PP_R element in " + file +
                               " to double");
    if (A->r.size() != (size_t)meshSize)
      throw Error(DFX_E_ARG, "Mismatch in number of radial points specified and the number of "
                             " radial points provided in " + file);
    DFX_REQUIRE(meshSize >= 2, "the radial mesh needs at least two points: " + file);
    for (size_t i = 0; i + 1 < A->r.size(); ++i)
      DFX_REQUIRE(A->r[i + 1] > A->r[i], "the radial mesh is not strictly increasing: " + file);
    const std::vector<double> &r = A->r;
    // Y_00 = C_00 D_0 Q_0 (SphericalHarmonicFunctions.cpp:378-430)
    const double invY00 = 1.0 / (std::sqrt(1.0 / 2.0) * (1.0 / std::sqrt(2 * M_PI)) * 1.0);
    static const std::map<std::string, std::string> elementOf = {
      {"vlocal", "PP_LOCAL"}, {"beta", "PP_NONLOCAL"}, {"pswfc", "PP_PSWFC"},
      {"nlcc", "PP_NLCC"},   {"rhoatom", "PP_RHOATOM"}};
    for (uint32_t fi = 0; fi < nFields; ++fi)
      {
        DFX_REQUIRE(fieldNames[fi], "null field name");
        const std::string fname = fieldNames[fi];
        const auto        el    = elementOf.find(fname);
        if (el == elementOf.end())
          throw Error(DFX_E_ARG, "No element name '" + fname +
                                   "' found in UPF_PARENT_PATH_LOOKUP. The fieldNames vector "
                                   "for UPF are standarized by authors.");
        const std::string xpath = "/UPF/" + el->second;
        std::vector<std::pair<const std::string *, std::vector<int>>> entries;
        std::vector<std::vector<int>>                                 seen;
        auto add = [&](const std::string *t, std::vector<int> q) {
          if (std::find(seen.begin(), seen.end(), q) != seen.end())
            throw Error(DFX_E_ARG, "Trying to enter more than one " + xpath +
                                     " of same quantum number vector from " + file);
          seen.push_back(q);
          entries.emplace_back(t, q);
        };
        auto attr = [](const XNode &n, const char *name) -> const std::string * {
          for (const auto &a : n.attrs)
            if (a.first == name)
              return &a.second;
          return nullptr;
        };
        static const std::string zeros;
        std::string zeroText;
        if (fname == "beta")
          {
            const XNode *nl = single(root, "PP_NONLOCAL", xpath);
            int l = 0, p = 0, lPrev = 0, lMaxBeta = 0;
            for (size_t i = 0; i < nl->kids.size(); ++i)
              {
                const XNode &c = nl->kids[i];
                if (c.name.find("PP_BETA") != std::string::npos)
                  {
                    const std::string *am = attr(c, "angular_momentum");
                    std::vector<int>   li;
                    if (!am || !toInts(*am, li) || li.size() != 1)
                      throw Error(DFX_E_ARG, "Error while converting angular_momentum in " +
                                               xpath + " element in " + file);
                    l = li[0];
                    if (l < 0)
                      throw Error(DFX_E_ARG, "Non-positive l found for" + xpath +
                                               " element in " + file);
                    // projector index within a run of equal l (the reference's
                    // counter: lPrev is first set at the second child)
                    if (i != 0)
                      {
                        p     = (l != lPrev) ? 0 : p + 1;
                        lPrev = l;
                      }
                    for (int m = -l; m <= l; ++m)
                      add(&c.text, {p, l, m});
                    lMaxBeta = std::max(lMaxBeta, l);
                  }
                else if (c.name.find("PP_DIJ") != std::string::npos)
                  {
                    std::vector<double> dij;
                    if (!toDoubles(c.text, dij) || dij.size() != (size_t)nProj * nProj)
                      throw Error(DFX_E_ARG, "Error while converting values in " + xpath +
                                               " element in " + file + " to double");
                    std::string out;
                    char        buf[64];
                    for (double v : dij)
                      {
                        snprintf(buf, sizeof buf, "%.16f ", v * 0.5); // Ry -> Ha
                        out += buf;
                      }
                    A->text["dij"] = {out};
                  }
              }
            if (lMaxBeta != lmax)
              throw Error(DFX_E_ARG, "The lmax from PSP HEADER does not match from" + xpath +
                                       " element in " + file);
          }
        else if (fname == "pswfc")
          {
            const XNode *pw = single(root, "PP_PSWFC", xpath);
            for (const XNode &c : pw->kids)
              if (c.name.find("PP_CHI") != std::string::npos)
                {
                  const std::string *ls = attr(c, "l"), *lab = attr(c, "label");
                  std::vector<int>   li;
                  if (!ls || !toInts(*ls, li) || li.size() != 1 || li[0] < 0 || !lab ||
                      lab->empty() || !isdigit((unsigned char)(*lab)[0]))
                    throw Error(DFX_E_ARG, "Error while converting l / label in " + xpath +
                                             " element in " + file);
                  const int l = li[0], n = (*lab)[0] - '0';
                  for (int m = -l; m <= l; ++m)
                    add(&c.text, {n, l, m});
                }
          }
        else if (fname == "nlcc" && !coreCorr)
          {
            for (int i = 0; i < meshSize; ++i)
              zeroText += i ? " 0" : "0";
            add(&zeroText, {0, 0, 0});
          }
        else
          add(&single(root, el->second.c_str(), xpath)->text, {0, 0, 0});

        std::vector<dfx_atom::Function> fun;
        for (const auto &e : entries)
          {
            dfx_atom::Function F;
            if (!toDoubles(*e.first, F.y))
              throw Error(DFX_E_ARG, "Error while converting values in " + xpath + " element in " +
                                       file + " to double");
            if (F.y.size() != (size_t)meshSize)
              throw Error(DFX_E_ARG, "Mismatch in number of points specified and number of points "
                                     "provided in " + xpath + " element in " + file);
            auto &v = F.y;
            if (fname == "vlocal")
              for (double &x : v)
                x *= 0.5;
            else if (fname == "beta" || fname == "pswfc")
              {
                for (size_t i = 1; i < v.size(); ++i)
                  v[i] = v[i] / r[i];
                v[0] = v[1];
              }
            else if (fname == "rhoatom")
              {
                for (size_t i = 1; i < v.size(); ++i)
                  v[i] = v[i] / (4 * M_PI * r[i] * r[i]);
                v[0] = v[1];
              }
            double cutoff = 1.0e10;
            for (size_t j = v.size() - 1; j > 0; --j)
              if (std::abs(v[j]) >= 1e-10)
                {
                  cutoff = j != v.size() - 1 ? r[j] : 1.0e10;
                  break;
                }
            if (fname == "rhoatom")
              cutoff = r.back();
            if (fname != "beta" && fname != "pswfc")
              for (double &x : v)
                x *= invY00;
            std::copy(e.second.begin(), e.second.end(), F.q);
            F.cutoff     = cutoff;
            F.smoothness = 1.0e10;
            if (fname == "vlocal")
              {
                std::vector<double> xs, ys;
                for (size_t i = 0; i < v.size(); ++i)
                  if (r[i] <= 10.0001)
                    {
                      xs.push_back(r[i]);
                      ys.push_back(v[i]);
                    }
                DFX_REQUIRE(xs.size() >= 2, "fewer than two vlocal points within 10 bohr: " + file);
                const double left  = (ys[1] - ys[0]) / (xs[1] - xs[0]);
                const double right = -1.0 * (ys.back() / xs.back());
                F.x                = xs;
                F.y                = ys;
                F.mixed            = true;
                F.tail             = -1.0 * invY00 * std::abs(zv);
                cubicSpline(F.x, F, true, left, false, right);
              }
            else
              cubicSpline(r, F);
            fun.push_back(std::move(F));
          }
        A->fields[fname] = std::move(fun);
      }
    if (!A->text.count("dij"))
      A->text["dij"] = {""};
    *out = A.release();
    return 0;
  });
}

int dfx_atom_function_knots(const dfx_atom *a, const char *field, uint32_t k, uint32_t *n,
                            double *knots, int *mixed, double *tail, double *cLeft)
{
  return guard(nullptr, [&] {
    const dfx_atom::Function &F  = functionOf(a, field, k);
    const std::vector<double> &x = F.knots(a->r);
    if (n)
      *n = (uint32_t)x.size();
    if (knots)
      std::copy(x.begin(), x.end(), knots);
    if (mixed)
      *mixed = F.mixed ? 1 : 0;
    if (tail)
      *tail = F.tail;
    if (cLeft)
      *cLeft = F.cLeft;
    return 0;
  });
}

int dfx_atom_metadata(const dfx_atom *a, const char *name, char *buf, uint64_t len,
                      uint64_t *needed)
{
  return guard(nullptr, [&] {
    DFX_REQUIRE(a && name, "null argument");
    const auto it = a->text.find(name);
    DFX_REQUIRE(it != a->text.end(), std::string("XPath expression /dft-efe:atom/dft-efe:") + name +
                                       " not found.");
    DFX_REQUIRE(it->second.size() == 1, std::string("Found more than one /dft-efe:atom/dft-efe:") +
                                          name + " element in " + a->file);
    const std::string v = trim(it->second[0]);
    if (needed)
      *needed = v.size() + 1;
    if (buf && len)
      {
        const size_t k = std::min<size_t>(v.size(), len - 1);
        memcpy(buf, v.data(), k);
        buf[k] = 0;
      }
    return 0;
  });
}

int dfx_atom_radial_grid(const dfx_atom *a, uint32_t *NR, double *r)
{
  return guard(nullptr, [&] {
    DFX_REQUIRE(a, "null atom data");
    if (NR)
      *NR = (uint32_t)a->r.size();
    if (r)
      std::copy(a->r.begin(), a->r.end(), r);
    return 0;
  });
}

int dfx_atom_num_functions(const dfx_atom *a, const char *field, uint32_t *count)
{
  return guard(nullptr, [&] {
    DFX_REQUIRE(a && field && count, "null argument");
    const auto it = a->fields.find(field);
    DFX_REQUIRE(it != a->fields.end(), std::string("field ") + field + " was not read from " + a->file);
    *count = (uint32_t)it->second.size();
    return 0;
  });
}

int dfx_atom_function(const dfx_atom *a, const char *field, uint32_t k, int *qNumbers,
                      double *cutoff, double *smoothness, double *values, double *splineB,
                      double *splineC, double *splineD)
{
  return guard(nullptr, [&] {
    const dfx_atom::Function &F = functionOf(a, field, k);
    if (qNumbers)
      std::copy(F.q, F.q + 3, qNumbers);
    if (cutoff)
      *cutoff = F.cutoff;
    if (smoothness)
      *smoothness = F.smoothness;
    if (values)
      std::copy(F.y.begin(), F.y.end(), values);
    if (splineB)
      std::copy(F.b.begin(), F.b.end(), splineB);
    if (splineC)
      std::copy(F.c.begin(), F.c.end(), splineC);
    if (splineD)
      std::copy(F.d.begin(), F.d.end(), splineD);
    return 0;
  });
}

int dfx_atom_evaluate(dfx_ctx *ctx, const dfx_atom *a, const char *field, uint32_t k,
                      const double *origin, dfx_mv *points, dfx_mv *values, dfx_mv *gradients)
{
  return guard(ctx, [&] {
    const dfx_atom::Function &F = functionOf(a, field, k);
    DFX_REQUIRE(ctx && origin && points && values, "null argument");
    DFX_REQUIRE(points->dtype == DFX_F64 && points->B == 3, "points: FP64, 3 columns (x, y, z)");
    DFX_REQUIRE(values->dtype == DFX_F64 && values->B == 1 && values->N == points->N,
                "values: FP64, one column, one row per point");
    DFX_REQUIRE(!gradients || (gradients->dtype == DFX_F64 && gradients->B == 3 &&
                               gradients->N == points->N),
                "gradients: FP64, 3 columns, one row per point");
    if (F.q[1] > 2)
      throw Error(DFX_E_UNSUPPORTED, "radial functions with l > 2 are not supported");
    const std::vector<double> &xs = F.knots(a->r);
    const size_t n  = xs.size();
    DevBuf      &sb = work(ctx, "atom.spline", 5 * n * sizeof(double));
    double      *d  = sb.as<double>();
    hipStream_t  s  = ctx->stream;
    const std::vector<double> *parts[5] = {&xs, &F.y, &F.b, &F.c, &F.d};
    for (int p = 0; p < 5; ++p)
      DFX_HIP(hipMemcpyAsync(d + p * n, parts[p]->data(), n * sizeof(double),
                             hipMemcpyHostToDevice, s));
    const SplineDev S{d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, (u32)n, F.cLeft};
    const u32       nPts = points->N;
    if (nPts)
      {
        atomEvalK<<<(nPts + 255) / 256, 256, 0, s>>>(
          S, F.q[1], F.q[2], F.cutoff, F.smoothness, F.mixed ? 1 : 0, F.tail, origin[0],
          origin[1], origin[2],
          static_cast<const double *>(points->d), nPts, static_cast<double *>(values->d),
          gradients ? static_cast<double *>(gradients->d) : nullptr);
        checkLaunch("atomEval");
      }
    // the host arrays of the copies stay owned by the atom data; wait so the
    // caller may destroy it on return
    DFX_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

int dfx_atom_destroy(dfx_atom *a)
{
  delete a;
  return 0;
}

} // extern "C"
