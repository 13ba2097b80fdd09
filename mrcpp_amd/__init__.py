"""mrcpp_amd: MI355X-native MRCPP operator-application path."""
