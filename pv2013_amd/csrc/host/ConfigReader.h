// Key/value configuration (ConfigReader.h:9-12 interface; grammar of the
// commented-out ConfigReader.cpp:13-96: "key value" lines, '#' comments,
// "Equation ... / Conditions:" blocks).
#pragma once

#include <map>
#include <string>
#include <utility>
#include <vector>

class ConfigReader {
public:
    explicit ConfigReader(const std::string& configFile);
    // Numeric values (std::stod).  Lines whose value is not a number are left
    // to readConfigStrings instead of failing the whole read.
    bool readConfigData(std::map<std::string, double>& constants);
    // Every "key value..." line; repeated keys keep all their values in order.
    bool readConfigStrings(std::multimap<std::string, std::string>& entries);
    bool readEquations(std::vector<std::pair<std::string, std::string>>& equations);
    void printEquations(const std::vector<std::pair<std::string, std::string>>& equations) const;
    const std::string& error() const { return err_; }

private:
    std::string configFile;
    std::string err_;
};
