#include "ini_file.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace tbdhost {

namespace {

const char* kWhite = " \f\n\r\t\v";

std::string TrimTwoEnds(const std::string& s) {
  const size_t a = s.find_first_not_of(kWhite);
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(kWhite);
  return s.substr(a, b - a + 1);
}

std::string TrimComment(const std::string& s) {
  const size_t p = s.find_first_of(";#");
  return p == std::string::npos ? s : s.substr(0, p);
}

// Recursive-descent evaluator: expr := term {(+|-) term}
// term := unary {(*|/) unary} ; unary := (+|-) unary | power
// power := primary [^ unary] (right associative) ; primary := number |
// constant | function(expr) | ( expr )
class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  double Parse() {
    const double v = Expr();
    Skip();
    if (i_ != s_.size()) Fail("unexpected '" + s_.substr(i_) + "'");
    return v;
  }

 private:
  void Fail(const std::string& m) const {
    throw std::runtime_error("IniFile::GetValueDouble(string,string) > Invalid Expression '" + s_ + "': " + m);
  }
  void Skip() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  bool Eat(char c) {
    Skip();
    if (i_ < s_.size() && s_[i_] == c) { ++i_; return true; }
    return false;
  }
  double Expr() {
    double v = Term();
    for (;;) {
      if (Eat('+')) v = v + Term();
      else if (Eat('-')) v = v - Term();
      else return v;
    }
  }
  double Term() {
    double v = Unary();
    for (;;) {
      if (Eat('*')) v = v * Unary();
      else if (Eat('/')) v = v / Unary();
      else return v;
    }
  }
  double Unary() {
    if (Eat('-')) return -Unary();
    if (Eat('+')) return Unary();
    return Power();
  }
  double Power() {
    const double b = Primary();
    if (Eat('^')) return std::pow(b, Unary());
    return b;
  }
  double Primary() {
    Skip();
    if (i_ >= s_.size()) Fail("unexpected end");
    if (Eat('(')) {
      const double v = Expr();
      if (!Eat(')')) Fail("missing ')'");
      return v;
    }
    const char c = s_[i_];
    if (std::isdigit((unsigned char)c) || c == '.') {
      size_t used = 0;
      double v = 0;
      try {
        v = std::stod(s_.substr(i_), &used);
      } catch (...) {
        Fail("bad number");
      }
      i_ += used;
      return v;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i_;
      while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_')) ++j;
      std::string id = s_.substr(i_, j - i_);
      i_ = j;
      std::string lid = id;
      std::transform(lid.begin(), lid.end(), lid.begin(), ::tolower);
      if (lid == "pi") return 3.14159265358979323846264338327950288;
      if (lid == "epsilon") return std::numeric_limits<double>::epsilon();
      if (lid == "inf") return std::numeric_limits<double>::infinity();
      if (Eat('(')) {
        const double a = Expr();
        if (!Eat(')')) Fail("missing ')'");
        if (lid == "sin") return std::sin(a);
        if (lid == "cos") return std::cos(a);
        if (lid == "tan") return std::tan(a);
        if (lid == "sqrt") return std::sqrt(a);
        if (lid == "exp") return std::exp(a);
        if (lid == "log") return std::log(a);
        if (lid == "abs") return std::fabs(a);
        Fail("unknown function '" + id + "'");
      }
      Fail("unknown symbol '" + id + "'");
    }
    Fail(std::string("unexpected '") + c + "'");
    return 0;
  }
  std::string s_;
  size_t i_ = 0;
};

std::vector<std::string> Tokens(const std::string& s) {
  std::istringstream in(s);
  return std::vector<std::string>(std::istream_iterator<std::string>(in), std::istream_iterator<std::string>());
}

}  // namespace

double EvaluateExpression(const std::string& text) { return Parser(text).Parse(); }

std::string UpperCase(const std::string& s) {
  std::string r = s;
  std::transform(r.begin(), r.end(), r.begin(), ::toupper);
  return r;
}

void IniFile::Load(const std::string& path) {
  std::ifstream in(path);
  if (!in.is_open()) throw std::runtime_error("IniFile::Load(string,ioctrl_t) > Can not Open the Configuration File!");
  sections_.clear();
  std::string line, section;
  while (std::getline(in, line)) {
    const size_t p = line.find_first_of(";#[=");
    if (p == std::string::npos) continue;
    if (line[p] == '[') {
      const size_t q = line.find_last_of(']');
      if (q == std::string::npos || q < p) continue;
      section = TrimTwoEnds(line.substr(p + 1, q - p - 1));
      bool found = false;
      for (auto& s : sections_) found = found || s.first == section;
      if (!found) sections_.push_back({section, {}});
    } else if (line[p] == '=') {
      const std::string key = TrimTwoEnds(line.substr(0, p));
      const std::string value = TrimTwoEnds(TrimComment(line.substr(p + 1)));
      for (auto& s : sections_)
        if (s.first == section) s.second.insert({key, value});   // first occurrence wins
    }
  }
}

bool IniFile::HasKey(const std::string& section, const std::string& key) const {
  for (auto& s : sections_)
    if (s.first == section) return s.second.count(key) > 0;
  return false;
}

std::string IniFile::GetValueString(const std::string& section, const std::string& key) const {
  for (auto& s : sections_) {
    if (s.first != section) continue;
    auto it = s.second.find(key);
    if (it == s.second.end())
      throw std::runtime_error("IniFile::GetValueString(string,string) > Can not Find Key: " + key + " in Section " +
                               section);
    return it->second;
  }
  throw std::runtime_error("IniFile::GetValueString(string,string) > Can not Find Section: " + section);
}

int IniFile::GetValueInt(const std::string& section, const std::string& key) const {
  const std::string v = GetValueString(section, key);
  try {
    return std::stoi(v);
  } catch (...) {
    throw std::runtime_error("IniFile::GetValueInt(string,string) > Invalid Integer for Key: " + key);
  }
}

bool IniFile::GetValueBool(const std::string& section, const std::string& key) const {
  return GetValueInt(section, key) != 0;
}

double IniFile::GetValueDouble(const std::string& section, const std::string& key) const {
  return EvaluateExpression(GetValueString(section, key));
}

std::vector<int> IniFile::GetVectorInt(const std::string& section, const std::string& key) const {
  std::vector<int> r;
  for (const auto& t : Tokens(GetValueString(section, key))) r.push_back((int)EvaluateExpression(t));
  return r;
}

std::vector<double> IniFile::GetVectorDouble(const std::string& section, const std::string& key) const {
  std::vector<double> r;
  for (const auto& t : Tokens(GetValueString(section, key))) r.push_back(EvaluateExpression(t));
  return r;
}

}  // namespace tbdhost
