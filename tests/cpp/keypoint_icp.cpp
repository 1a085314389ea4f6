// The reference's keyPointICP (RealTimeRobot/function.h:111-127) written
// against include/rtr_registration.hpp: load two clouds, register the model
// onto the scan, print the transform, transform the model and save it.
// Usage: keypoint_icp <source.pcd> <target.pcd> <out.pcd> [--io-only]
#include <cstring>
#include <iostream>

#include "rtr_registration.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: keypoint_icp source.pcd target.pcd out.pcd [--io-only]\n";
    return 2;
  }
  auto mcloud = std::make_shared<rtr::PointCloud<rtr::PointXYZ>>();
  auto cloud = std::make_shared<rtr::PointCloud<rtr::PointXYZ>>();
  if (rtr::io::loadPCDFile(argv[1], *mcloud) != 0 || rtr::io::loadPCDFile(argv[2], *cloud) != 0) {
    std::cerr << "load failed: " << rtr_last_error() << "\n";
    return 1;
  }
  std::cout << "source " << mcloud->size() << " target " << cloud->size() << "\n";
  if (argc > 4 && std::strcmp(argv[4], "--io-only") == 0) {
    return rtr::io::savePCDFileASCII(argv[3], *mcloud) == 0 ? 0 : 1;
  }
  try {
    rtr::IterativeClosestPoint<rtr::PointXYZ, rtr::PointXYZ> icp;
    icp.setInputCloud(mcloud);
    icp.setInputTarget(cloud);
    icp.setMaximumIterations(50);
    rtr::PointCloud<rtr::PointXYZ> Final;
    icp.align(Final);
    std::cout << icp.getFinalTransformation() << std::endl;
    std::cout << "converged " << icp.hasConverged() << " fitness " << icp.getFitnessScore() << "\n";
    rtr::PointCloud<rtr::PointXYZ> transformed;
    rtr::transformPointCloud(*mcloud, transformed, icp.getFinalTransformation());
    for (size_t i = 0; i < transformed.size(); ++i)
      if (transformed[i].x != Final[i].x || transformed[i].y != Final[i].y || transformed[i].z != Final[i].z) {
        std::cerr << "align output and transformPointCloud differ at " << i << "\n";
        return 1;
      }
    return rtr::io::savePCDFileASCII(argv[3], transformed) == 0 ? 0 : 1;
  } catch (const rtr::Error& e) {
    std::cerr << e.what() << "\n";
    return 3;
  }
}
