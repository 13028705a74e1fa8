// Key/value configuration (ConfigReader.h:9-10 interface, ConfigReader.cpp:13-51
// grammar): "key value" lines, '#' comments.
#pragma once

#include <map>
#include <string>

class ConfigReader {
public:
    explicit ConfigReader(const std::string& configFile);
    // Numeric values (std::stod).  Lines whose value is not a number are left
    // to readConfigStrings instead of failing the whole read.
    bool readConfigData(std::map<std::string, double>& constants);
    // Every "key value..." line; repeated keys keep all their values in order.
    bool readConfigStrings(std::multimap<std::string, std::string>& entries);
    const std::string& error() const { return err_; }

private:
    std::string configFile;
    std::string err_;
};
