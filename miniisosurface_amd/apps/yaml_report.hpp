/*
 * yaml_report.hpp -- the miniapp drivers' YAML run report, same layout and file naming.
 *
 * Layout (mantevoCommon YAML_Doc / YAML_Element as used by flyingEdges/serial/main.cpp:90-205):
 *   Mini-Application Name: <name>
 *   Mini-Application Version: <version>
 *   <key>: <value>                 values printed with std::ostream defaults
 *     <child key>: <child value>   two more spaces per level
 * File: with a -y path "dir/file", the report goes to dir + "/" + file (the directory is created);
 * without one, to ./<name>-<version>_<YYYY:MM:DD-HH:MM:SS>.yaml. The text is also returned for
 * printing to standard output.
 */
#ifndef MINIISO_APPS_YAML_REPORT_HPP
#define MINIISO_APPS_YAML_REPORT_HPP

#include <sys/stat.h>

#include <ctime>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace miso_apps {

class YamlNode {
public:
    YamlNode(std::string key, std::string value) : key_(std::move(key)), value_(std::move(value)) {}

    template <typename V>
    YamlNode* add(const std::string& key, const V& value)
    {
        std::ostringstream s;
        s << value;
        children_.push_back(std::make_unique<YamlNode>(key, s.str()));
        return children_.back().get();
    }
    YamlNode* add(const std::string& key, const char* value) { return add(key, std::string(value)); }

    YamlNode* get(const std::string& key)
    {
        for (auto& c : children_)
            if (c->key_ == key) return c.get();
        return nullptr;
    }

    std::string print(const std::string& indent) const
    {
        std::string out = indent + key_ + ": " + value_ + "\n";
        for (const auto& c : children_) out += c->print(indent + "  ");
        return out;
    }

protected:
    std::string key_, value_;
    std::vector<std::unique_ptr<YamlNode> > children_;
    friend class YamlReport;
};

class YamlReport : public YamlNode {
public:
    /* yaml_arg: the -y argument, or "" when none was given */
    YamlReport(std::string name, std::string version, const std::string& yaml_arg)
        : YamlNode("", ""), name_(std::move(name)), version_(std::move(version))
    {
        if (!yaml_arg.empty()) {
            const size_t pos = yaml_arg.rfind('/');
            if (pos == std::string::npos) {
                dir_ = "./";
                file_ = yaml_arg;
            } else {
                dir_ = yaml_arg.substr(0, pos + 1);
                file_ = yaml_arg.substr(pos + 1);
            }
        }
    }

    /* writes the report file and returns its text */
    std::string generate()
    {
        std::string out = "Mini-Application Name: " + name_ + "\n";
        out += "Mini-Application Version: " + version_ + "\n";
        for (const auto& c : children_) out += c->print("");
        std::string path;
        if (!dir_.empty() && dir_ != ".") {
            ::mkdir(dir_.c_str(), 0755);
            path = dir_ + "/" + file_;
        } else {
            std::time_t now = std::time(nullptr);
            std::tm tmv{};
            localtime_r(&now, &tmv);
            char date[80];
            std::snprintf(date, sizeof(date), "%04d:%02d:%02d-%02d:%02d:%02d", tmv.tm_year + 1900, tmv.tm_mon + 1,
                          tmv.tm_mday, tmv.tm_hour, tmv.tm_min, tmv.tm_sec);
            path = "./" + (file_.empty() ? name_ + "-" + version_ + "_" : file_) + date + ".yaml";
        }
        std::ofstream f(path);
        f << out;
        return out;
    }

private:
    std::string name_, version_, dir_, file_;
};

}  // namespace miso_apps

#endif
