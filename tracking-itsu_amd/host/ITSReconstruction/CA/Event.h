// Event.h -- input containers of the reference API: Cluster, Layer, Event
// (include/ITSReconstruction/CA/{Cluster,Layer,Event}.h; Event.cxx:26-68,
// Layer.cxx:38-42, Cluster.cxx:27-33).
#pragma once

#include <array>
#include <vector>

#include "Constants.h"
#include "Definitions.h"

namespace o2 {
namespace ITS {
namespace CA {

// Same fields as the reference Cluster (Cluster.h:36-44).  Raw clusters hold
// zero phi, r and bin; the tracker computes them on the device.
struct Cluster {
  Cluster(int clusterId, int layerIndex, float x, float y, float z, float alpha, int mc)
      : xCoordinate{x}, yCoordinate{y}, zCoordinate{z}, phiCoordinate{0}, rCoordinate{0}, clusterId{clusterId},
        alphaAngle{alpha}, monteCarloId{mc}, indexTableBinIndex{0}
  {
    (void)layerIndex;
  }
  Cluster() = default;

  float xCoordinate;
  float yCoordinate;
  float zCoordinate;
  float phiCoordinate;
  float rCoordinate;
  int clusterId;
  float alphaAngle;
  int monteCarloId;
  int indexTableBinIndex;
};

class Layer {
 public:
  Layer() : mLayerIndex{Constants::ITS::UnusedIndex} {}
  explicit Layer(int layerIndex) : mLayerIndex{layerIndex} {}
  int getLayerIndex() const { return mLayerIndex; }
  const std::vector<Cluster>& getClusters() const { return mClusters; }
  const Cluster& getCluster(int i) const { return mClusters[i]; }
  int getClustersSize() const { return (int)mClusters.size(); }
  void addCluster(int clusterId, float x, float y, float z, float alpha, int mc)
  {
    mClusters.emplace_back(clusterId, mLayerIndex, x, y, z, alpha, mc);
  }

 private:
  int mLayerIndex;
  std::vector<Cluster> mClusters;
};

class Event {
 public:
  explicit Event(int eventId) : mEventId{eventId}
  {
    for (int l = 0; l < Constants::ITS::LayersNumber; ++l) mLayers[l] = Layer(l);
  }
  int getEventId() const { return mEventId; }
  const float3& getPrimaryVertex(int i) const { return mPrimaryVertices[i]; }
  const Layer& getLayer(int l) const { return mLayers[l]; }
  int getPrimaryVerticesNum() const { return (int)mPrimaryVertices.size(); }
  void addPrimaryVertex(float x, float y, float z) { mPrimaryVertices.push_back(float3{x, y, z}); }
  void pushClusterToLayer(int layer, int clusterId, float x, float y, float z, float alpha, int mc)
  {
    mLayers[layer].addCluster(clusterId, x, y, z, alpha, mc);
  }
  int getTotalClusters() const
  {
    int n = 0;
    for (const Layer& l : mLayers) n += l.getClustersSize();
    return n;
  }
  void printPrimaryVertices() const;

 private:
  int mEventId;
  std::vector<float3> mPrimaryVertices;
  std::array<Layer, Constants::ITS::LayersNumber> mLayers;
};

}  // namespace CA
}  // namespace ITS
}  // namespace o2
