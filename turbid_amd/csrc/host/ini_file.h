// INI configuration reader with the reference's semantics
// (reference src/file_ini.h:22-110, src/file_ini.cc:43-293):
//  - a line's first of ";#[=" decides: '[' opens a section, '=' a key,
//    ';' / '#' a comment; text after ';' or '#' in a value is dropped;
//  - keys and section names are case sensitive;
//  - integers parse with std::stoi, doubles and vector entries are
//    evaluated as arithmetic expressions (exprtk in the reference; here the
//    subset the configuration files use: + - * / ^, parentheses, unary
//    signs, decimal/exponent literals and the constants pi, epsilon, inf);
//  - a missing section or key throws
//    "IniFile::GetValueString(string,string) > Can not Find ...".
#pragma once

#include <map>
#include <string>
#include <vector>

namespace tbdhost {

double EvaluateExpression(const std::string& text);

class IniFile {
 public:
  void Load(const std::string& path);
  bool HasKey(const std::string& section, const std::string& key) const;
  std::string GetValueString(const std::string& section, const std::string& key) const;
  int GetValueInt(const std::string& section, const std::string& key) const;
  bool GetValueBool(const std::string& section, const std::string& key) const;
  double GetValueDouble(const std::string& section, const std::string& key) const;
  std::vector<int> GetVectorInt(const std::string& section, const std::string& key) const;
  std::vector<double> GetVectorDouble(const std::string& section, const std::string& key) const;

 private:
  std::vector<std::pair<std::string, std::map<std::string, std::string>>> sections_;
};

// ConvertCase(target, 1): upper case (reference src/file_ini.h:44-52)
std::string UpperCase(const std::string& s);

}  // namespace tbdhost
