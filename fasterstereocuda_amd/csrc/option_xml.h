// Option files of the stereo console: the XML schema of the reference's
// FasterStereoConsole/option_manager.cpp:46-208 (loading) and :210-340
// (saving), read and written by a small XML reader of this project.
//
//   <Option1>                         disparity form (StereoOption1)
//     <View>   width height </View>
//     <Stereo> disp_min disp_max layers_num do_consistency_check
//              do_remove_peaks do_smooth </Stereo>
//   <Option2>                         depth form (StereoOption2)
//     <View>   width height </View>
//     <Stereo> depth_min depth_max layers_num do_consistency_check
//              do_remove_peaks do_smooth </Stereo>
//     <Camera> x0_left_pix y0_left_pix x0_right_pix y0_right_pix
//              baseline_mm focus_pix </Camera>
//
// Every element is optional; a missing one leaves the option's default.
// Values are read as C atoi / atof read them (leading blanks skipped,
// trailing text ignored, no digits -> 0).
#pragma once

#include <string>
#include <vector>

#include "FasterStereoCuda.h"

namespace fsc {

// One element of a parsed document: its name, its character data (entities
// decoded, surrounding blanks kept) and its child elements.
struct XmlElement {
    std::string name;
    std::string text;
    std::vector<XmlElement> children;
    const XmlElement* child(const char* n) const;   // first child of that name
};

// Parses a whole document; returns false (and a message) on malformed XML.
bool parse_xml(const std::string& doc, XmlElement* root, std::string* err);

enum class OptionKind { kDisparity, kDepth, kNone };

// OptionManager::GaussType (option_manager.cpp:17-43): the root element's
// name decides the form; kNone when the file is missing, unreadable or has
// another root.
OptionKind option_kind(const std::string& path);

// OptionManager::Load (option_manager.cpp:46-208).  A missing file is
// written with the current values of `opt` and the call returns false, as in
// the reference.  A file that does not parse returns false and is left as it
// is (the reference overwrites it with the defaults).
bool load_option(const std::string& path, FasterStereoCuda::ste_opt1_t& opt);
bool load_option(const std::string& path, FasterStereoCuda::ste_opt2_t& opt);

// OptionManager::Save (option_manager.cpp:210-340): the declaration, then
// one element per field, four-space indents, numbers as std::to_string
// prints them.
bool save_option(const std::string& path, const FasterStereoCuda::ste_opt1_t& opt);
bool save_option(const std::string& path, const FasterStereoCuda::ste_opt2_t& opt);

}  // namespace fsc
